"""CPU: the C-ABI library and the drop-in modules load, export what the
header declares, and validate arguments the way pptk's extensions do
(pptk/kdtree/kdtree_wrapper.cpp:281-412,
pptk/processing/estimate_normals/estimate_normals.cpp:259-345) before any
device work."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "pptk_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pptk_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from pptk_amd import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(capi.EXPORTED) == syms


def test_modules_import_with_reference_names():
    import pptk_amd
    assert pptk_amd.kdtree.__name__ == "pptk_amd.kdtree.kdtree"
    assert callable(pptk_amd.kdtree._build) and callable(pptk_amd.kdtree._query)
    assert pptk_amd.estimate_normals.__name__ == "estimate_normals"
    assert pptk_amd.kdtree.__file__.startswith(os.path.join(ROOT, "pptk_amd"))


def test_last_error_empty_and_device_count():
    from pptk_amd import capi
    assert capi.lib().pptk_last_error() == b""
    assert capi.lib().pptk_device_count() >= 0


def test_query_argument_errors():
    from pptk_amd import kdtree
    with pytest.raises(TypeError):
        kdtree._query(object(), None, k=1)
    with pytest.raises(ValueError):
        kdtree._query(object(), None, k=0)
    with pytest.raises(ValueError):
        kdtree._query(object(), None, k=-3)
    with pytest.raises(ValueError):
        kdtree._query(object(), None, k="a")       # int('a')
    with pytest.raises(ValueError):
        kdtree._query(object(), None, dmax=-1.0)
    with pytest.raises(ValueError):
        kdtree._query(object(), None, dmax="a")    # float('a')
    with pytest.raises(TypeError):
        kdtree._query()


def test_build_argument_errors():
    from pptk_amd import kdtree
    with pytest.raises(RuntimeError):
        kdtree._build(np.zeros((10, 5)))            # dim 5
    with pytest.raises(RuntimeError):
        kdtree._build(np.zeros(10))                 # one 10-D point
    with pytest.raises(RuntimeError):
        kdtree._build(np.zeros((10, 3), np.complex64))
    with pytest.raises(ValueError):
        kdtree._build(np.zeros((2, 2, 3)))          # ndim 3
    with pytest.raises(TypeError):
        kdtree._build("points")
    with pytest.raises(TypeError):
        kdtree._build(np.zeros((4, 3)), maxleafsize=2.5)


def test_estimate_normals_argument_errors():
    from pptk_amd import estimate_normals
    y = np.zeros((10, 3))
    with pytest.raises(ValueError):
        estimate_normals(y, 4, 0.0)
    with pytest.raises(ValueError):
        estimate_normals(y, 0, 1.0)
    with pytest.raises(ValueError):
        estimate_normals(y, -1, np.inf)
    with pytest.raises(RuntimeError):
        estimate_normals(y, "a", 1.0)
    with pytest.raises(TypeError):
        estimate_normals(y.astype(np.int32), 4, 1.0)
    with pytest.raises(ValueError):
        estimate_normals(y, 4, 1.0, subsample=np.zeros(2, dtype=bool))
    with pytest.raises(RuntimeError):
        estimate_normals(y, 4, 1.0, subsample=[0, 10])
    with pytest.raises(TypeError):
        estimate_normals(y, 4, 1.0, subsample="x")
    with pytest.raises(TypeError):
        estimate_normals("x", 4, 1.0)
    with pytest.raises(ValueError):
        estimate_normals(y, 4, 1.0, precision="half")


def test_no_device_fails_loudly():
    from pptk_amd import capi, kdtree
    if capi.lib().pptk_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no HIP device"):
        kdtree._build(np.random.rand(10, 3))
    with pytest.raises(RuntimeError, match="no HIP device"):
        import pptk_amd
        pptk_amd.estimate_normals(np.random.rand(10, 3), 3, 1.0, verbose=False)


def test_capi_host_checks_without_device():
    from pptk_amd import capi
    L = capi.lib()
    h = ctypes.c_void_p()
    bp = capi.BuildParams(0, -1, 0.2)
    st = L.pptk_tree_build(None, 10, 5, capi.F32, ctypes.byref(bp), capi.MEM_HOST, None,
                           ctypes.byref(h))
    assert st == capi.E_RUNTIME and b"dim" in L.pptk_last_error()
    st = L.pptk_tree_build(None, 10, 3, 77, ctypes.byref(bp), capi.MEM_HOST, None,
                           ctypes.byref(h))
    assert st == capi.E_RUNTIME
    assert L.pptk_tree_check(None) == 0
    assert L.pptk_tree_check(ctypes.c_void_p(12345)) == 0
    q = capi.Queries()
    st = L.pptk_knn(ctypes.c_void_p(12345), ctypes.byref(q), 1, 1.0, None, None, None, 0, None)
    assert st == capi.E_TYPE


def test_scenes_deterministic():
    from pptk_amd import scenes
    a = scenes.lidar_scene(50_000, seed=1)
    b = scenes.lidar_scene(50_000, seed=1)
    assert a.dtype == np.float32 and a.shape == (50_000, 3) and np.array_equal(a, b)
    r = np.sqrt((a.astype(np.float64) ** 2).sum(1))
    # 1/range^2 density: about half the points within ~13 m of the sensor
    assert 5 < np.median(r) < 25
    u = scenes.uniform_cube(1000, seed=0)
    np.random.seed(0)
    assert np.array_equal(u, np.random.rand(1000, 3))
    cl = scenes.batched_clouds(count=5, seed=3)
    assert len(cl) == 5 and all(c.dtype == np.float32 and c.shape[1] == 3 for c in cl)

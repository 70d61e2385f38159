"""CPU tests: the C-ABI library loads and exports every symbol include/gsp.h
declares; the host-side mirror of the reference interface (grids, initialization,
method defaults, ensembles) behaves like the reference's tests expect."""
import os
import re

import numpy as np
import pytest

import geostatsprocesses as gsp
from geostatsprocesses import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols(header="gsp.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|const char\*)\s+(gsp_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 16
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES)
    testing = declared_symbols("gsp_testing.h")
    assert testing and set(testing) == set(_lib.TESTING_SIGNATURES)
    for s in testing:
        assert hasattr(lib, s), s
    assert lib.gsp_abi_version() == 1


def test_no_device_is_a_clean_error():
    """Without a GPU the context call fails with a status code, never a crash."""
    import ctypes
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.gsp_ctx_create(0, ctypes.byref(h))
    if rc == 0:  # a GPU is visible
        lib.gsp_ctx_destroy(h)
        pytest.skip("GPU present")
    assert rc == _lib.GSP_ERR_DEVICE
    assert b"device" in lib.gsp_last_error()


def test_null_arguments_are_rejected():
    import ctypes
    lib = _lib.load()
    assert lib.gsp_ctx_create(0, None) == _lib.GSP_ERR_ARGUMENT
    assert lib.gsp_seqsim_plan_info(None, None) == _lib.GSP_ERR_ARGUMENT
    assert lib.gsp_lusim_dims(None, None, None) == _lib.GSP_ERR_ARGUMENT
    assert lib.gsp_seqsim_sweep(None, None, 1, None, None, None, 0) == _lib.GSP_ERR_ARGUMENT
    out = ctypes.c_void_p()
    assert lib.gsp_seqsim_plan_create(None, None, ctypes.byref(out)) == _lib.GSP_ERR_ARGUMENT


# ------------------------------------------------------------- host logic ----
def load_tsv(name, ncoord):
    arr = np.loadtxt(os.path.join(GOLDEN, name), skiprows=1)
    return gsp.georef({"value": arr[:, ncoord]}, [tuple(r) for r in arr[:, :ncoord]])


def test_grid_linear_indices_column_major():
    g = gsp.CartesianGrid(100, 100)
    assert g.linear_index(24, 24) == 2424
    assert g.linear_index(1, 0) == 1 and g.linear_index(0, 1) == 100
    np.testing.assert_allclose(g.centroid(2424), [24.5, 24.5])
    g = gsp.CartesianGrid((0.5, 0.5), (100.5, 100.5), dims=(100, 100))
    np.testing.assert_allclose(g.centroid(g.linear_index(24, 24)), [25.0, 25.0])


def test_nearest_init_reference_mappings():
    """test/initialization.jl:5-29 (1-based indices there)."""
    proc = gsp.GaussianProcess(gsp.GaussianVariogram())
    init = gsp.NearestInit()
    data1D = load_tsv("data1D.tsv", 1)
    grid = gsp.CartesianGrid((-0.5,), (99.5,), dims=(100,))
    real, mask = gsp.initialize(proc, grid, data1D, init)
    for i, j in {100: 11, 81: 9, 11: 2, 21: 3, 91: 10, 51: 6, 61: 7, 71: 8, 31: 4, 41: 5, 1: 1}.items():
        assert real["value"][i - 1] == data1D.value[j - 1]
        assert mask["value"][i - 1]
    data2D = load_tsv("data2D.tsv", 2)
    grid = gsp.CartesianGrid((-0.5, -0.5), (99.5, 99.5), dims=(100, 100))
    real, mask = gsp.initialize(proc, grid, data2D, init)
    for i, j in {5076: 3, 2526: 1, 7551: 2}.items():
        assert real["value"][i - 1] == data2D.value[j - 1]
        assert mask["value"][i - 1]
    pset = gsp.PointSet([(25, 25), (50, 75), (75, 50)])
    real, mask = gsp.initialize(proc, pset, data2D, init)
    for i, j in {2: 2, 3: 3, 1: 1}.items():
        assert real["value"][i - 1] == data2D.value[j - 1]
        assert mask["value"][i - 1]


def test_explicit_init_reference_placements():
    """test/initialization.jl:31-59."""
    proc = gsp.GaussianProcess(gsp.GaussianVariogram())
    grid = gsp.CartesianGrid(10, 10, 10)
    rng = np.random.default_rng(0)
    data = gsp.georef({"z": rng.random(10)}, [tuple(p) for p in rng.random((10, 3))])
    real, mask = gsp.initialize(proc, grid, data, gsp.ExplicitInit(np.arange(10)))
    assert np.array_equal(real["z"][:10], data.z) and mask["z"][:10].all()
    real, mask = gsp.initialize(proc, grid, data, gsp.ExplicitInit(np.arange(990, 1000)))
    assert np.array_equal(real["z"][990:], data.z) and mask["z"][990:].all()
    real, mask = gsp.initialize(proc, grid, data, gsp.ExplicitInit(np.arange(3), np.arange(997, 1000)))
    assert np.array_equal(real["z"][997:], data.z[:3]) and mask["z"].sum() == 3
    with pytest.raises(AssertionError):
        gsp.initialize(proc, grid, data, gsp.ExplicitInit(np.arange(3), np.arange(5)))


def test_defaultsimulation_reference_choices():
    """test/field.jl:3-15."""
    proc1 = gsp.GaussianProcess(gsp.GaussianVariogram())
    proc2 = gsp.GaussianProcess(gsp.GaussianCovariance())
    grid = gsp.CartesianGrid(100, 100)
    rng = np.random.default_rng(1)
    pset1 = gsp.PointSet(rng.random((1000, 2)))
    pset2 = gsp.PointSet(rng.random((10000, 2)))
    assert isinstance(gsp.defaultsimulation(proc1, grid), gsp.FFTSIM)
    assert isinstance(gsp.defaultsimulation(proc1, pset1), gsp.SEQSIM)
    assert isinstance(gsp.defaultsimulation(proc1, pset2), gsp.SEQSIM)
    assert isinstance(gsp.defaultsimulation(proc2, pset1), gsp.LUSIM)


def test_process_and_method_defaults():
    proc = gsp.GaussianProcess(gsp.SphericalVariogram(range=35.0))
    assert proc.mean == 0.0 and proc.defaultnames() == ("field",)
    m = gsp.SEQSIM()
    assert (m.minneighbors, m.maxneighbors, m.neighborhood) == (1, 26, ":range")
    assert isinstance(m.path, gsp.LinearPath)
    with pytest.raises(AssertionError):
        gsp.GaussianProcess(gsp.SphericalVariogram(), [0.0, 1.0])
    f = gsp.ExponentialVariogram(ranges=(40.0, 40.0, 8.0), nugget=0.05)
    assert f.range() == 40.0 and not f.isbanded()
    assert gsp.ExponentialCovariance(range=20.0).isbanded()


def test_paths_are_permutations():
    g = gsp.CartesianGrid(7, 5)
    for p in (gsp.LinearPath(), gsp.RandomPath(seed=3), gsp.SourcePath([4, 2, 4])):
        order = p.traverse(g)
        assert sorted(order) == list(range(35))
    assert list(gsp.SourcePath([4, 2, 4]).traverse(g)[:3]) == [4, 2, 0]


def test_async_on_master_is_argument_error():
    """test/field.jl:321: async with the master among the workers."""
    proc = gsp.GaussianProcess(gsp.GaussianVariogram())
    with pytest.raises(gsp.ArgumentError):
        gsp.rand(proc, gsp.CartesianGrid(10, 10), 3, workers=[0], async_=True)


def test_ensemble_container():
    """ensembles.jl:20-35: length, indexing and iteration (the reductions run
    on the GPU, tests/test_ensemble_gpu.py)."""
    g = gsp.CartesianGrid(3)
    vals = np.array([[0.0, 1.0, 2.0], [2.0, 3.0, 4.0], [1.0, 2.0, 9.0]])
    e = gsp.Ensemble(g, {"field": vals})
    assert len(e) == 3
    assert np.array_equal(e[1].field, vals[1])
    assert [np.array_equal(t.field, vals[i]) for i, t in enumerate(e)] == [True] * 3
    assert repr(e) == "1D Ensemble"


@pytest.mark.parametrize("cls,fam", [(gsp.SphericalTransiogram, 0), (gsp.ExponentialTransiogram, 1),
                                     (gsp.GaussianTransiogram, 2)])
def test_transiogram_is_a_coregionalization(cls, fam):
    """Shape transiograms: T(0) = I, rows of T(h) sum to one, T(h) = 1 p^T at
    and beyond the range (spherical) or tends to it, and the indicator
    cross-covariances p_i (T_ij(h) - p_j) equal c(h) (diag(p) - p p^T) with c
    the oracle's unit-sill covariance of the family (1 - v(h)): the reduction
    that lets SEQSIM simulate an IndicatorProcess of a transiogram with the
    simple-kriging weights of c."""
    import oracle as O
    p = np.array([0.7, 0.2, 0.1])
    t = cls(range=35.0, proportions=tuple(p))
    np.testing.assert_array_equal(t.transition(0.0), np.eye(3))
    A = np.diag(p) - np.outer(p, p)
    for h in (0.5, 3.0, 17.5, 34.0, 60.0, 200.0):
        T = t.transition(h)
        np.testing.assert_allclose(T.sum(axis=1), 1.0, rtol=0, atol=1e-15)
        c = O.cov_u(fam, 1.0, 0.0, h / 35.0)
        np.testing.assert_allclose(p[:, None] * (T - p[None, :]), c * A, rtol=1e-13, atol=1e-16)
    if fam == 0:
        np.testing.assert_allclose(t.transition(35.0), np.broadcast_to(p, (3, 3)), atol=1e-15)
    proc = gsp.IndicatorProcess(t)
    np.testing.assert_allclose(proc.prob, p)
    assert proc.func.nvariables() == 3
    # the default IndicatorProcess(transiogram) proportions are two equal ones
    assert np.allclose(gsp.IndicatorProcess(cls(range=10.0)).prob, [0.5, 0.5])

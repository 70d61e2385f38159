"""Pins the CPU oracle (oracle/gsp_oracle.c, oracle/oracle.py) before it is
trusted as the checker: against closed forms, independent numpy/scipy
computations, brute force on small inputs and the reference's own test
expectations. CPU only."""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O


# ------------------------------------------------------------------- exp ----
def test_exp_matches_numpy_within_two_ulp():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-745, 709, 20000), rng.uniform(-1, 1, 5000),
                         [0.0, -0.0, 1e-300, -1e-300, 0.5 * np.log(2), -0.5 * np.log(2),
                          -3.0, -30.0, -700.0, -744.0, 709.0]])
    ours = np.array([O.exp(x) for x in xs])
    ref = np.exp(xs)
    fin = ref > 0
    ulp = np.abs(ours[fin] - ref[fin]) / np.spacing(ref[fin])
    assert ulp.max() <= 2.0
    assert O.exp(0.0) == 1.0
    assert O.exp(-800.0) == 0.0 and np.isinf(O.exp(800.0))


# ------------------------------------------------------------ variograms ----
@pytest.mark.parametrize("fam", [O.SPHERICAL, O.EXPONENTIAL, O.GAUSSIAN])
def test_covariance_invariants(fam):
    sill, nug = 2.0, 0.3
    assert O.cov_u(fam, sill, nug, 0.0) == sill           # gamma(0) = 0
    us = np.linspace(1e-6, 3, 400)
    c = np.array([O.cov_u(fam, sill, nug, u) for u in us])
    assert np.all(np.diff(c) <= 1e-15)                     # monotone non-increasing
    assert abs(c[0] - (sill - nug)) < 1e-5                 # nugget jump at 0+
    assert c.min() >= -1e-15


def test_covariance_closed_forms():
    # spherical: C(u) = s (1 - 3u/2 + u^3/2) for u < 1 (nugget 0), 0 beyond
    for u in (0.1, 0.25, 0.5, 0.9):
        assert np.isclose(O.cov_u(O.SPHERICAL, 1.5, 0.0, u), 1.5 * (1 - 1.5 * u + 0.5 * u ** 3),
                          rtol=1e-14)
    assert O.cov_u(O.SPHERICAL, 1.0, 0.0, 1.0) == 0.0
    assert O.cov_u(O.SPHERICAL, 1.0, 0.0, 7.0) == 0.0
    # practical range: 95% of the structured sill is reached at u = 1
    assert np.isclose(O.cov_u(O.EXPONENTIAL, 1.0, 0.0, 1.0), np.exp(-3.0), rtol=1e-14)
    assert np.isclose(O.cov_u(O.EXPONENTIAL, 1.0, 0.01, 0.5), 0.99 * np.exp(-1.5), rtol=1e-14)
    # gaussian: the structured part is sill - nugget - 1e-6 (stabilising nugget)
    assert np.isclose(O.cov_u(O.GAUSSIAN, 1.0, 0.0, 1.0), (1.0 - 1e-6) * np.exp(-3.0), rtol=1e-14)
    assert np.isclose(O.cov_u(O.GAUSSIAN, 1.0, 0.0, 0.5), (1.0 - 1e-6) * np.exp(-0.75), rtol=1e-14)
    assert np.isclose(O.cov_u(O.GAUSSIAN, 2.0, 0.5, 0.5), (1.5 - 1e-6) * np.exp(-0.75), rtol=1e-14)
    assert O.cov_u(O.GAUSSIAN, 2.0, 0.5, 0.0) == 2.0


def _centroids_1d(n):
    return (np.arange(n, dtype=np.float64) + 0.5)[:, None]


def test_gaussian_reference_lusim_cases_are_positive_definite():
    """The reference's LUSIM testset factors GaussianCovariance matrices of unit
    grids (test/field.jl:29-37: range 10 on CartesianGrid(100); :46-56: range
    35 on a 10 x 10 grid, with and without data) and expects fields back. With
    the plain Gaussian kernel those matrices are singular in FP64 (LAPACK's
    potrf fails); with the stabilising nugget they factor, the smallest
    eigenvalue then being at least that nugget."""
    import scipy.linalg as sla
    cases = [(_centroids_1d(100), (10.0,)),
             (O.grid_centroids((10, 10)), (35.0, 35.0))]
    for cents, ranges in cases:
        C = O.covariance_np(O.GAUSSIAN, 1.0, 0.0, ranges, cents, cents)
        L = sla.cholesky(C, lower=True)
        assert np.all(np.isfinite(L))
        assert np.linalg.eigvalsh(C).min() > 0.5e-6
        # the same kernel without the stabilising nugget is not positive definite
        u = np.sqrt(np.sum(((cents[:, None, :] - cents[None, :, :]) / np.asarray(ranges)) ** 2, -1))
        with pytest.raises(np.linalg.LinAlgError):
            sla.cholesky(np.exp(-3.0 * u * u), lower=True)
    # the conditional 10 x 10 case (data at (25,25), (50,75), (75,50) snap to cells)
    cents = O.grid_centroids((10, 10))
    dinds = np.unique(O.nearest_init(cents, np.array([[25.0, 25.0], [50.0, 75.0], [75.0, 50.0]])))
    d2, L22, _ = O.lusim_preprocess(cents, dinds, np.zeros(len(dinds)), O.GAUSSIAN, 1.0, 0.0,
                                    (35.0, 35.0))
    assert np.all(np.isfinite(L22))


# --------------------------------------------------------------- scaling ----
def test_data_only_kriging_in_domain_coordinates():
    """GeoStatsModels.fitpredict (posterior mean, FFTSIM conditioning) kriges in
    the domain's own coordinates: alpha = 1 for data-only plans, so data on cell
    centroids sit at lag exactly 0 and are reproduced exactly even where the
    model has a nugget (a gaussian model always has its stabilising one). With
    SEQSIM's scaling by 1/100.5 the lag would round to a tiny positive value and
    the nugget jump would move the estimate by about 1e-6."""
    coords = np.array([[25.0, 25.0], [50.0, 75.0], [75.0, 50.0]])
    vals = np.array([1.0, 0.0, 1.0])
    case = O.SeqsimCase(dims=(100, 100), origin=(0.5, 0.5), family=O.GAUSSIAN, ranges=(35.0,),
                        data_coords=coords, data_vals=vals, kmax=3, mode=O.DATA_ONLY)
    assert O.scale_factor(case) == 1.0
    plan = O.seqsim_plan(case)
    mean = O.seqsim_sweep(case, plan, np.zeros((1, case.n_nodes)))[0]
    for (x, y), v in zip(coords, vals):
        i = int(x - 1.0) + 100 * int(y - 1.0)
        assert mean[i] == v


def test_scale_factor_closed_form():
    # seqsim.jl:162-189: alpha = 1 / max(|grid bbox|, |data bbox|, range)
    c = O.SeqsimCase(dims=(100, 100), ranges=(10.0,))
    assert O.scale_factor(c) == 1.0 / 100.0
    c = O.SeqsimCase(dims=(100, 100), ranges=(250.0,))
    assert O.scale_factor(c) == 1.0 / 250.0
    c = O.SeqsimCase(dims=(10, 10), origin=(-40.0, 0.0), ranges=(5.0,),
                     data_coords=np.array([[3.0, -70.5]]))
    assert O.scale_factor(c) == 1.0 / 70.5


def test_clamp_neighbors():
    # seqsim.jl:202-208
    assert O.clamp_neighbors(100, 1, 26) == (1, 26)
    assert O.clamp_neighbors(10, 1, 26) == (1, 10)
    assert O.clamp_neighbors(10, 5, 0) == (5, 10)
    assert O.clamp_neighbors(10, 12, 8) == (1, 8)


# ---------------------------------------------------------------- search ----
def numpy_search(case, target, pos):
    """Independent restatement: sort every augmented point by (d2, index)."""
    N = case.n_nodes
    cents = O.grid_centroids(case.dims, case.origin, case.spacing)
    tgt_ijk = np.array(np.unravel_index(target, case.dims, order="F"))
    all_ijk = np.array(np.unravel_index(np.arange(N), case.dims, order="F")).T
    d2g = np.sum(((all_ijk - tgt_ijk) * np.asarray(case.spacing)) ** 2, axis=1)
    d2d = np.sum((case.data_coords - cents[target]) ** 2, axis=1)
    d2 = np.concatenate([d2g, d2d])
    idx = np.arange(N + case.n_data)
    elig = np.concatenate([pos < pos[target], np.ones(case.n_data, bool)])
    inball = d2 <= case.radius ** 2 if case.radius > 0 else np.ones_like(d2, bool)
    _, k = O.clamp_neighbors(N, case.kmin, case.kmax)
    order = np.lexsort((idx, d2))
    order = order[inball[order]]
    if case.mode == O.DATA_ONLY:
        return order[order >= N][:k].astype(np.int32)
    if case.mode == O.KNN_THEN_MASK:
        top = order[:k]
        return top[elig[top]].astype(np.int32)
    order = order[elig[order]]
    return order[:k].astype(np.int32)


CASES = [
    dict(dims=(40,), ranges=(6.0,), kmax=5, radius=0.0, nd=4),
    dict(dims=(25, 18), ranges=(8.0,), kmax=12, radius=6.0, nd=12),
    dict(dims=(25, 18), ranges=(8.0,), kmax=16, radius=0.0, nd=0),
    dict(dims=(9, 8, 7), ranges=(4.0, 4.0, 2.0), kmax=24, radius=4.0, nd=10),
    dict(dims=(12, 10), spacing=(1.0, 2.5), ranges=(8.0,), kmax=8, radius=7.5, nd=30),
]


@pytest.mark.parametrize("mode", [O.KNN_THEN_MASK, O.MASKED_KNN, O.DATA_ONLY])
@pytest.mark.parametrize("ci", range(len(CASES)))
def test_search_brute_and_bucket_match_numpy(ci, mode):
    spec = dict(CASES[ci])
    rng = np.random.default_rng(ci)
    nd = spec.pop("nd")
    if mode == O.DATA_ONLY and nd == 0:
        nd = 7
    dims = spec["dims"]
    N = int(np.prod(dims))
    sp = np.asarray(spec.get("spacing", (1.0,) * len(dims)))
    # some data outside the grid box, some exactly on node centres (ties)
    data = rng.uniform(-0.1, 1.1, (nd, len(dims))) * np.asarray(dims) * sp
    if nd >= 2:
        data[0] = (np.array([3] + [2] * (len(dims) - 1)) + 0.5) * sp
    if mode == O.MASKED_KNN and spec["radius"] == 0.0:
        spec["radius"] = 6.0
    case = O.SeqsimCase(**spec, data_coords=data, path=rng.permutation(N), mode=mode)
    pos = O.positions(case)
    for t in range(N):
        ref = numpy_search(case, t, pos)
        assert np.array_equal(O.search(case, t, brute=True, pos=pos), ref), t
        assert np.array_equal(O.search(case, t, brute=False, pos=pos), ref), t


def test_search_ties_lower_index_wins():
    # 5x5 grid, target in the centre, all 4 axial neighbours at distance 1
    first = [17, 13, 11, 7, 12]            # all four axial neighbours precede 12
    path = np.array(first + [i for i in range(25) if i not in first])
    case = O.SeqsimCase(dims=(5, 5), kmax=3, radius=0.0, mode=O.MASKED_KNN, path=path)
    got = O.search(case, 12)
    assert list(got) == [7, 11, 13]  # the four at d=1 are 7, 11, 13, 17: lowest three


def test_search_knn_then_mask_counts_target():
    # k nearest of ALL points include the target itself, which is then masked
    first = [17, 13, 11, 7, 12]
    path = np.array(first + [i for i in range(25) if i not in first])
    case = O.SeqsimCase(dims=(5, 5), kmax=3, radius=0.0, mode=O.KNN_THEN_MASK, path=path)
    got = O.search(case, 12)
    assert list(got) == [7, 11]


# --------------------------------------------------------------- kriging ----
def numpy_krige(case, target, nbr):
    alpha = O.scale_factor(case)
    N = case.n_nodes
    cents = O.grid_centroids(case.dims, case.origin, case.spacing)
    pts = np.array([cents[a] if a < N else case.data_coords[a - N] for a in nbr])
    rng_ = case.ranges[: case.ndim]
    LHS = O.covariance_np(case.family, case.sill, case.nugget, rng_, pts, pts, case.rotation)
    c0 = O.covariance_np(case.family, case.sill, case.nugget, rng_, pts, cents[target:target + 1],
                         case.rotation)[:, 0]
    lam = np.linalg.solve(LHS, c0)
    var = max(case.sill - lam @ c0, 0.0)
    return lam, np.sqrt(var), alpha


@pytest.mark.parametrize("fam", [O.SPHERICAL, O.EXPONENTIAL, O.GAUSSIAN])
def test_krige_matches_numpy_solve(fam):
    rng = np.random.default_rng(7)
    case = O.SeqsimCase(dims=(30, 20, 6), family=fam, ranges=(9.0, 7.0, 3.0), nugget=0.1,
                        kmax=20, radius=9.0, data_coords=rng.uniform(0, 1, (25, 3)) * [30, 20, 6],
                        path=rng.permutation(3600))
    pos = O.positions(case)
    checked = 0
    for t in rng.choice(3600, 60, replace=False):
        nbr = O.search(case, t, pos=pos)
        if len(nbr) < 2:
            continue
        lam, sig, f = O.krige(case, t, nbr)
        lam_np, sig_np, _ = numpy_krige(case, t, nbr)
        assert f == 0
        assert np.allclose(lam, lam_np, rtol=1e-9, atol=1e-12)
        assert np.isclose(sig, sig_np, rtol=1e-9, atol=1e-12)
        checked += 1
    assert checked > 30


def test_krige_rotation_matches_numpy():
    th = 0.4
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    case = O.SeqsimCase(dims=(20, 20), family=O.EXPONENTIAL, ranges=(10.0, 3.0), nugget=0.05,
                        rotation=R, kmax=16, radius=10.0)
    for t in (45, 210, 399):
        nbr = O.search(case, t)
        if len(nbr) < 2:
            continue
        lam, sig, f = O.krige(case, t, nbr)
        lam_np, sig_np, _ = numpy_krige(case, t, nbr)
        assert np.allclose(lam, lam_np, rtol=1e-9, atol=1e-12)
        assert np.isclose(sig, sig_np, rtol=1e-9)


def test_krige_single_neighbour_closed_form():
    # k = 1: lambda = C(h)/C(0), var = C(0) - C(h)^2/C(0)
    case = O.SeqsimCase(dims=(10,), family=O.EXPONENTIAL, ranges=(4.0,), sill=2.0, nugget=0.2)
    lam, sig, f = O.krige(case, 5, np.array([3]))
    ch = O.cov_u(O.EXPONENTIAL, 2.0, 0.2, 2.0 / 4.0)
    assert np.isclose(lam[0], ch / 2.0, rtol=1e-14)
    assert np.isclose(sig ** 2, 2.0 - ch * ch / 2.0, rtol=1e-13)


def test_krige_colocated_datum_honours_data():
    # a datum on the target's centre with nugget 0: lambda = e_j, sigma = 0
    case = O.SeqsimCase(dims=(10, 10), family=O.SPHERICAL, ranges=(5.0,),
                        data_coords=np.array([[4.5, 4.5]]))
    N = 100
    lam, sig, f = O.krige(case, 44, np.array([N, 43, 34]))
    assert f == 0
    assert np.allclose(lam, [1.0, 0.0, 0.0], atol=1e-12)
    assert sig < 1e-6


def test_krige_singular_system_fails_to_prior():
    # two data on the same spot, nugget 0: LHS is singular -> status false
    # (a gaussian model would not do: its stabilising nugget keeps it regular)
    case = O.SeqsimCase(dims=(10, 10), family=O.SPHERICAL, ranges=(50.0,),
                        data_coords=np.array([[2.2, 2.2], [2.2, 2.2]]))
    lam, sig, f = O.krige(case, 0, np.array([100, 101]))
    assert f & O.FLAG_FAIL
    assert np.all(lam == 0) and sig == 1.0


def test_fp32_refine_close_to_fp64():
    rng = np.random.default_rng(3)
    case = O.SeqsimCase(dims=(24, 24, 8), family=O.EXPONENTIAL, ranges=(10.0, 10.0, 4.0),
                        nugget=0.05, kmax=24, radius=10.0, path=rng.permutation(24 * 24 * 8))
    pos = O.positions(case)
    errs = []
    for t in rng.choice(case.n_nodes, 80, replace=False):
        nbr = O.search(case, t, pos=pos)
        if len(nbr) < 2:
            continue
        l64, s64, _ = O.krige(case, t, nbr, precision=O.FP64)
        l32, s32, f = O.krige(case, t, nbr, precision=O.FP32_REFINE)
        errs.append(np.max(np.abs(l64 - l32)))
        assert abs(s64 - s32) <= 1e-5
    assert max(errs) <= 1e-5


# ----------------------------------------------------------------- sweep ----
def test_plan_sweep_equals_reference_cost_structure():
    """Hoisting search + solve out of the realization loop changes nothing."""
    rng = np.random.default_rng(11)
    case = O.SeqsimCase(dims=(30, 25), family=O.SPHERICAL, ranges=(8.0,), kmax=12, radius=15.0,
                        data_coords=rng.uniform(0, 25, (20, 2)), data_vals=rng.standard_normal(20),
                        path=rng.permutation(750))
    z = rng.standard_normal((3, 750))
    plan = O.seqsim_plan(case)
    a = O.seqsim_sweep(case, plan, z)
    b, n = O.seqsim_reference_cost(case, z, threads=2)
    assert n == 3 * 750
    assert np.array_equal(a, b)


def test_plan_levels_are_longest_chains():
    rng = np.random.default_rng(5)
    case = O.SeqsimCase(dims=(20, 20), ranges=(6.0,), kmax=8, radius=6.0, path=rng.permutation(400))
    p = O.seqsim_plan(case)
    for i in range(400):
        k = 0 if p.flags[i] & 3 else p.nfound[i]
        deps = [a for a in p.nbr[i, :k] if a < 400]
        assert p.level[i] == (1 + max(p.level[deps]) if deps else 0)


def test_seqsim_data_honouring_reference_case():
    """test/field.jl:73-86: SEQSIM(maxneighbors=3), SphericalVariogram(range=35),
    grid (0.5,0.5)-(100.5,100.5) 100x100, data at (25,25),(50,75),(75,50);
    the node at (25,25) reproduces its datum (atol 1e-4)."""
    rng = np.random.default_rng(2017)
    case = O.SeqsimCase(dims=(100, 100), origin=(0.5, 0.5), family=O.SPHERICAL, ranges=(35.0,),
                        kmax=3, radius=35.0, data_coords=np.array([[25.0, 25.0], [50.0, 75.0], [75.0, 50.0]]),
                        data_vals=np.array([1.0, 0.0, 1.0]), path=rng.permutation(10000))
    plan = O.seqsim_plan(case)
    out = O.seqsim_sweep(case, plan, rng.standard_normal((2, 10000)))
    i = 24 + 100 * 24  # inds[25, 25] (1-based) -> 0-based linear index
    assert np.allclose(out[:, i], 1.0, atol=1e-4)
    assert np.allclose(out[:, 74 + 100 * 49], 1.0, atol=1e-4)   # (75, 50)
    assert np.allclose(out[:, 49 + 100 * 74], 0.0, atol=1e-4)        # (50, 75)


def test_seqsim_unconditional_statistics():
    """Ensemble mean ~ mu and variance ~ sill for an unconditional simulation."""
    rng = np.random.default_rng(8)
    case = O.SeqsimCase(dims=(40, 40), family=O.EXPONENTIAL, ranges=(6.0,), kmax=16, radius=12.0,
                        mean=0.5, sill=2.0, path=rng.permutation(1600), mode=O.MASKED_KNN)
    plan = O.seqsim_plan(case)
    out = O.seqsim_sweep(case, plan, rng.standard_normal((200, 1600)))
    assert abs(out.mean() - 0.5) < 0.1
    assert abs(out.var(axis=0).mean() - 2.0) < 0.2


# ----------------------------------------------------------------- LUSIM ----
def test_lusim_oracle_matches_conditional_gaussian():
    """d2 = C21 C11^-1 z1 and L22 L22^T = C22 - C21 C11^-1 C12 (inverse-based)."""
    rng = np.random.default_rng(2)
    dims = (12, 10)
    cents = O.grid_centroids(dims)
    dinds = np.sort(rng.choice(120, 9, replace=False))
    z1 = rng.standard_normal(9)
    d2, L22, sinds = O.lusim_preprocess(cents, dinds, z1, O.EXPONENTIAL, 1.0, 0.01, (6.0, 6.0))
    C = O.covariance_np(O.EXPONENTIAL, 1.0, 0.01, (6.0, 6.0), cents, cents)
    C11, C12, C22 = C[np.ix_(dinds, dinds)], C[np.ix_(dinds, sinds)], C[np.ix_(sinds, sinds)]
    inv = np.linalg.inv(C11)
    assert np.allclose(d2, C12.T @ inv @ z1, rtol=1e-9, atol=1e-12)
    assert np.allclose(L22 @ L22.T, C22 - C12.T @ inv @ C12, atol=1e-11)
    out = O.lusim_sample(d2, L22, dinds, sinds, z1, 0.3, rng.standard_normal((4, len(sinds))))
    assert np.array_equal(out[:, dinds], np.broadcast_to(z1, (4, 9)))   # data reproduced exactly


def test_lusim_oracle_unconditional_adds_mean():
    cents = O.grid_centroids((8, 8))
    d2, L22, sinds = O.lusim_preprocess(cents, [], [], O.SPHERICAL, 1.0, 0.0, (4.0, 4.0))
    w = np.random.default_rng(1).standard_normal((2, 64))
    out = O.lusim_sample(d2, L22, np.zeros(0, np.int64), sinds, [], 0.7, w)
    assert np.allclose(out, 0.7 + w @ L22.T)


def test_nearest_init_reference_fixtures():
    """test/initialization.jl:9-28 expected cell mappings (1-based there)."""
    d1 = np.loadtxt(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                               "data1D.tsv"), skiprows=1)
    cents = O.grid_centroids((100,), origin=(-0.5,))
    dest = O.nearest_init(cents, d1[:, :1])
    expect = {100: 11, 81: 9, 11: 2, 21: 3, 91: 10, 51: 6, 61: 7, 71: 8, 31: 4, 41: 5, 1: 1}
    for i, j in expect.items():
        assert dest[j - 1] == i - 1
    d2 = np.loadtxt(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                               "data2D.tsv"), skiprows=1)
    cents = O.grid_centroids((100, 100), origin=(-0.5, -0.5))
    dest = O.nearest_init(cents, d2[:, :2])
    for i, j in {5076: 3, 2526: 1, 7551: 2}.items():
        assert dest[j - 1] == i - 1


# -------------------------------------------------------- posterior mean ----
def test_posterior_mean_reference_case():
    """test/field.jl:163-170: the simple kriging posterior mean reproduces the
    data at their cells (isapprox with Julia's default tolerance)."""
    case = O.SeqsimCase(dims=(100, 100), origin=(0.5, 0.5), family=O.SPHERICAL, ranges=(35.0,),
                        data_coords=np.array([[25.0, 25.0], [50.0, 75.0], [75.0, 50.0]]),
                        data_vals=np.array([1.0, 0.0, 1.0]), kmax=10, mode=O.DATA_ONLY)
    plan = O.seqsim_plan(case)
    m = O.seqsim_sweep(case, plan, np.zeros((1, case.n_nodes)))[0]
    for (i, j), v in {(24, 24): 1.0, (49, 74): 0.0, (74, 49): 1.0}.items():
        assert m[i + 100 * j] == pytest.approx(v, rel=np.sqrt(np.finfo(float).eps), abs=0.0)
    ref, var, _ = O.simple_kriging_np(case)
    np.testing.assert_allclose(m, ref, rtol=1e-10, atol=1e-12)
    assert np.all(plan.level == 0)


def test_simple_kriging_np_single_datum_closed_form():
    """One datum: lambda = C(h) / C(0), mean = mu + lambda (z - mu),
    variance = C(0) - C(h)^2 / C(0)."""
    x0 = np.array([[3.3, 4.1]])
    case = O.SeqsimCase(dims=(8, 6), family=O.EXPONENTIAL, sill=2.0, nugget=0.3, ranges=(5.0,),
                        mean=0.4, data_coords=x0, data_vals=np.array([1.7]), kmax=1,
                        mode=O.DATA_ONLY)
    m, v, ids = O.simple_kriging_np(case)
    X = O.grid_centroids(case.dims)
    h = np.linalg.norm(X - x0, axis=1)
    ch = (2.0 - 0.3) * np.exp(-3.0 * h / 5.0)
    np.testing.assert_allclose(m, 0.4 + ch / 2.0 * (1.7 - 0.4), rtol=1e-13)
    np.testing.assert_allclose(v, 2.0 - ch ** 2 / 2.0, rtol=1e-13)
    plan = O.seqsim_plan(case)
    mo = O.seqsim_sweep(case, plan, np.zeros((1, case.n_nodes)))[0]
    np.testing.assert_allclose(mo, m, rtol=1e-12)
    np.testing.assert_allclose(plan.sigma ** 2, v, rtol=1e-12)


@pytest.mark.parametrize("fam,kmax,radius", [(O.SPHERICAL, 10, 0.0), (O.GAUSSIAN, 6, 0.0),
                                             (O.EXPONENTIAL, 12, 4.0)])
def test_posterior_mean_oracle_matches_numpy(fam, kmax, radius):
    rng = np.random.default_rng(kmax)
    dims = (14, 11, 5)
    data = rng.uniform(-0.05, 1.05, (40, 3)) * np.asarray(dims)
    case = O.SeqsimCase(dims=dims, family=fam, sill=1.5, nugget=0.05, ranges=(6.0, 5.0, 3.0),
                        mean=0.2, data_coords=data, data_vals=rng.standard_normal(40),
                        kmin=2, kmax=kmax, radius=radius, mode=O.DATA_ONLY)
    plan = O.seqsim_plan(case)
    ref, var, ids = O.simple_kriging_np(case)
    N = case.n_nodes
    few = np.isnan(ref)
    assert radius == 0.0 or few.any()
    np.testing.assert_array_equal((plan.flags & 1) != 0, few)
    for i in np.flatnonzero(~few):
        n = plan.nfound[i]
        np.testing.assert_array_equal(plan.nbr[i, :n] - N, ids[i, :n])
    m = O.seqsim_sweep(case, plan, np.zeros((1, N)))[0]
    np.testing.assert_allclose(m[~few], ref[~few], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(plan.sigma[~few] ** 2, var[~few], rtol=1e-9, atol=1e-12)
    assert np.all(plan.level == 0)


# ---------------------------------------------------------------- FFTSIM ----
def circular_autocorrelation(x, shape):
    X = np.fft.fftn(x.reshape(shape))
    return np.real(np.fft.ifftn(X * np.conj(X)))


def test_fftsim_oracle_reproduces_covariance():
    """FFT moving average: every realization has |fft(Z)| = F exactly, so its
    circular autocorrelation is ifft(F^2) = ifft(|fft(C_shifted)|); for a
    covariance whose sampled spectrum is non-negative that is C itself, less
    its mean over the periodic window (F[1] = 0 removes the field's mean):
    A(h) / A(0) = (C(h) - Cbar) / (C(0) - Cbar) in every realization."""
    dims = (100, 80)
    F = O.fftsim_amplitude(dims, O.GAUSSIAN, 1.0, 0.0, (10.0, 10.0))
    w = np.random.default_rng(0).random((2, 8000))
    out = O.fftsim_unconditional(F, w, 1.0, 0.3)
    # the periodic window holds every lag once: -(n/2 - 1) .. n/2 per axis
    hx = np.arange(-49, 51)[None, :]
    hy = np.arange(-39, 41)[:, None]
    # the model: (1 - eps) exp(-3 h^2 / r^2) for h > 0 and 1 at h = 0, eps the
    # gaussian model's stabilising nugget (1e-6)
    eps = 1e-6
    Cw = (1.0 - eps) * np.exp(-3.0 * (hx ** 2 + hy ** 2) / 100.0)
    Cw[(hy == 0) & (hx == 0)] = 1.0
    cbar = np.mean(Cw)
    for r in range(2):
        A = circular_autocorrelation(out[r] - 0.3, F.shape)
        for (dy, dx) in [(0, 1), (1, 0), (2, 3), (5, 0), (0, 12), (7, 7)]:
            c = (1.0 - eps) * np.exp(-3.0 * (dx * dx + dy * dy) / 100.0)
            expect = (c - cbar) / (1.0 - cbar)
            assert A[dy, dx] / A[0, 0] == pytest.approx(expect, abs=1e-9)
    # a dropped square root would give the covariance convolved with itself
    A2 = np.real(np.fft.ifftn(F ** 4))
    assert abs(A2[0, 5] / A2[0, 0] - np.exp(-3.0 * 0.25)) > 0.05


def test_fftsim_oracle_invariants():
    """var(Z - mean) = sill exactly; the result depends on the noise only
    through its phases: invariant to scaling the noise and, as F[1] = 0, to
    adding a constant to it."""
    dims = (24, 18, 10)
    F = O.fftsim_amplitude(dims, O.EXPONENTIAL, 2.0, 0.1, (6.0, 5.0, 3.0))
    assert F.flat[0] == 0.0 and F.shape == (10, 18, 24)
    w = np.random.default_rng(1).random((1, F.size))
    out = O.fftsim_unconditional(F, w, 2.0, -1.0)[0]
    assert np.sum((out + 1.0) ** 2) / (out.size - 1) == pytest.approx(2.0, rel=1e-12)
    np.testing.assert_allclose(O.fftsim_unconditional(F, 3.7 * w, 2.0, -1.0)[0], out, atol=1e-12)
    np.testing.assert_allclose(O.fftsim_unconditional(F, w + 0.5, 2.0, -1.0)[0], out, atol=1e-12)


def test_fftsim_oracle_conditional_honours_centred_data():
    """test/field.jl:134-141 setting: data on cell centroids are reproduced
    (zbar honours the data and the kriged unconditional values equal the
    unconditional values at the data cells)."""
    dims = (100, 100)
    coords = np.array([[25.0, 25.0], [50.0, 75.0], [75.0, 50.0]])
    vals = np.array([1.0, 0.0, 1.0])
    origin = (0.5, 0.5)
    F = O.fftsim_amplitude(dims, O.GAUSSIAN, 1.0, 0.0, (35.0, 35.0), origin=origin)
    zu = O.fftsim_unconditional(F, np.random.default_rng(2).random((2, 10000)), 1.0, 0.0)
    dcase = O.SeqsimCase(dims=dims, origin=origin, family=O.GAUSSIAN, ranges=(35.0,),
                         data_coords=coords, data_vals=vals, kmax=3, mode=O.DATA_ONLY)
    zbar, _, _ = O.simple_kriging_np(dcase)
    cells = np.sort(O.nearest_init(O.grid_centroids(dims, origin), coords))
    ccase = O.SeqsimCase(dims=dims, origin=origin, family=O.GAUSSIAN, ranges=(35.0,),
                         data_coords=O.grid_centroids(dims, origin)[cells], kmax=3,
                         mode=O.DATA_ONLY)
    z = O.fftsim_conditional(zu, zbar, cells, O.simple_kriging_weights(ccase), 0.0)
    for c, v in zip(O.nearest_init(O.grid_centroids(dims, origin), coords), vals):
        np.testing.assert_allclose(z[:, c], v, atol=1e-9)
    assert not np.allclose(z[0], z[1])


# ---------------------------------------------------- ensemble statistics ----
def test_ensemble_stat_reference_testset():
    """test/ensembles.jl:24-59: three constant realizations 1, 2, 3."""
    reals = np.stack([i * np.ones(9) for i in (1.0, 2.0, 3.0)])
    assert np.all(O.ensemble_stat(reals, "mean") == 2.0)
    assert np.all(O.ensemble_stat(reals, "var") == 1.0)
    for i in (1, 2, 3):
        assert np.all(O.ensemble_stat(reals, "cdf", i) == i / 3)
        np.testing.assert_allclose(O.ensemble_stat(reals, "ccdf", i), 1 - i / 3)
    assert np.all(O.ensemble_stat(reals, "quantile", 0.5) == 2.0)
    assert np.all(O.ensemble_stat(reals, "quantile", 0.0) == 1.0)
    assert np.all(O.ensemble_stat(reals, "quantile", 1.0) == 3.0)


@pytest.mark.parametrize("R", [1, 2, 7, 16, 33])
def test_ensemble_stat_matches_numpy(R):
    """Independent routes: numpy's pairwise mean, two-pass var, and the
    'linear' (type 7) quantile agree to rounding; p = 0 gives the minimum
    exactly (g = 0), p = 1 the second largest plus the gap to the largest
    (g = 1), which is the maximum up to rounding."""
    v = np.random.default_rng(R).standard_normal((R, 500)) * 3.0 + 1.0
    np.testing.assert_allclose(O.ensemble_stat(v, "mean"), v.mean(axis=0), rtol=1e-13, atol=1e-14)
    if R > 1:
        np.testing.assert_allclose(O.ensemble_stat(v, "var"), v.var(axis=0, ddof=1), rtol=1e-12)
    else:
        assert np.all(np.isnan(O.ensemble_stat(v, "var")))
    for p in (0.0, 0.1, 0.25, 0.5, 0.9, 1.0):
        np.testing.assert_allclose(O.ensemble_stat(v, "quantile", p), np.quantile(v, p, axis=0),
                                   rtol=1e-13, atol=1e-14)
    assert np.array_equal(O.ensemble_stat(v, "quantile", 0.0), v.min(axis=0))
    if R > 1:
        srt = np.sort(v, axis=0)
        assert np.array_equal(O.ensemble_stat(v, "quantile", 1.0), srt[-2] + (srt[-1] - srt[-2]))
    np.testing.assert_allclose(O.ensemble_stat(v, "quantile", 1.0), v.max(axis=0), rtol=0,
                               atol=1e-15 * np.abs(v).max())
    np.testing.assert_array_equal(O.ensemble_stat(v, "cdf", 1.0), (v <= 1.0).mean(axis=0))
    np.testing.assert_array_equal(O.ensemble_stat(v, "ccdf", 1.0) + O.ensemble_stat(v, "cdf", 1.0),
                                  np.ones(500))


# ------------------------------------------------------- indicator SEQSIM ----
def test_cokriging_of_coregionalized_indicators_reduces_to_kriging():
    """Simple cokriging with C_ab(h) = A_ab c(h) and every variable observed at
    every neighbour: the full (k n) x (k n) system has the weights
    lambda_j I, lambda the simple-kriging weights of c(h) (the reduction the
    indicator sweep relies on)."""
    rng = np.random.default_rng(4)
    X = rng.uniform(0, 10, (7, 2))
    x0 = np.array([[4.0, 5.0]])
    A = np.array([[1.0, 0.4, 0.1], [0.4, 1.0, 0.3], [0.1, 0.3, 1.0]])
    C = O.covariance_np(O.SPHERICAL, 1.0, 0.05, (6.0, 6.0), X, X)
    c0 = O.covariance_np(O.SPHERICAL, 1.0, 0.05, (6.0, 6.0), X, x0)[:, 0]
    lam = np.linalg.solve(C, c0)
    Lam = np.linalg.solve(np.kron(C, A), np.kron(c0[:, None], A))  # (7*3) x 3
    for j in range(7):
        np.testing.assert_allclose(Lam[3 * j:3 * j + 3], lam[j] * np.eye(3), atol=1e-12)


def test_categorical_draw_matches_cumulative_search():
    """The first category whose running sum exceeds u, independently by
    searchsorted on the cumulative sums (the last category at most)."""
    rng = np.random.default_rng(5)
    q = rng.dirichlet(np.ones(4), 5000)
    u = rng.random(5000)
    u[:10] = 0.0
    u[10:20] = np.nextafter(1.0, 0.0)
    ref = np.minimum([np.searchsorted(np.cumsum(qq), uu, side="right") for qq, uu in zip(q, u)], 3)
    np.testing.assert_array_equal(O.draw_category(q, u), ref)
    # normalisation: clamped and scaled to unit sum
    raw = rng.normal(0.3, 0.4, (100, 3))
    raw[:, 0] = np.abs(raw[:, 0]) + 0.01  # every row keeps some mass
    out = O.indicator_probabilities(raw)
    assert np.all(out >= 0) and np.allclose(out.sum(1), 1.0)
    np.testing.assert_allclose(out, np.clip(raw, 0, 1) / np.clip(raw, 0, 1).sum(1, keepdims=True),
                               rtol=1e-15)


def test_indicator_sweep_prior_only_follows_probabilities():
    """With no neighbour inside the search ball every node draws from the
    prior: the category frequencies follow the probabilities."""
    case = O.SeqsimCase(dims=(40, 40), family=O.SPHERICAL, ranges=(8.0,), kmax=12, radius=0.5,
                        path=np.random.default_rng(1).permutation(1600))
    plan = O.seqsim_plan(case)
    assert np.all((plan.flags & 1) != 0)
    u = np.random.default_rng(2).random((4, 1600))
    out = O.indicator_sweep(case, plan, u, [0.6, 0.3, 0.1], [])
    freq = np.bincount(out.ravel(), minlength=3) / out.size
    np.testing.assert_allclose(freq, [0.6, 0.3, 0.1], atol=0.02)


# ------------------------------------------------- multivariate SEQSIM ----
def test_mv_factor_is_cholesky():
    A = np.array([[2.0, 0.9, -0.3], [0.9, 1.5, 0.4], [-0.3, 0.4, 1.1]])
    L = O.mv_factor(A)
    np.testing.assert_allclose(L, np.linalg.cholesky(A), rtol=1e-15, atol=1e-15)
    assert np.all(np.triu(L, 1) == 0)
    with pytest.raises(np.linalg.LinAlgError):
        O.mv_factor(np.array([[1.0, 2.0], [2.0, 1.0]]))


def test_multivariate_sweep_one_variable_is_the_scalar_sweep():
    rng = np.random.default_rng(31)
    case = O.SeqsimCase(dims=(14, 11), family=O.SPHERICAL, ranges=(6.0,), kmax=8, radius=7.0,
                        data_coords=rng.uniform(0, 11, (9, 2)), data_vals=rng.standard_normal(9),
                        mean=0.4, path=rng.permutation(154))
    plan = O.seqsim_plan(case)
    z = rng.standard_normal((3, 154))
    mv = O.seqsim_sweep_multivariate(case, plan, z[:, :, None], [[1.0]], [0.4],
                                     case.data_vals[None, :])
    np.testing.assert_array_equal(mv[0], O.seqsim_sweep(case, plan, z))


def test_multivariate_sweep_covariance_is_A_times_scalar():
    """Over many realizations the two fields of A * c(h) (A = [1 .8; .8 1])
    have, at every pair of nodes, cross-covariance A12 / A11 times the
    auto-covariance, the MvNormal semantics with covariance A sigma^2 (the
    transposed factor would give 0.62 instead of 0.8)."""
    rng = np.random.default_rng(32)
    case = O.SeqsimCase(dims=(9, 8), family=O.EXPONENTIAL, ranges=(4.0,), kmax=6, radius=0.0,
                        path=rng.permutation(72))
    plan = O.seqsim_plan(case)
    A = np.array([[1.0, 0.8], [0.8, 1.0]])
    R = 6000
    z = rng.standard_normal((R, 72, 2))
    x = O.seqsim_sweep_multivariate(case, plan, z, A, [0.0, 0.0], np.zeros((2, 0)))
    c11 = x[0].T @ x[0] / R
    c12 = x[0].T @ x[1] / R
    c22 = x[1].T @ x[1] / R
    # lag 0 correlation at every node and the ratio of cross- to auto-covariance
    corr = np.diag(c12) / np.sqrt(np.diag(c11) * np.diag(c22))
    assert abs(corr.mean() - 0.8) < 0.01
    big = np.abs(c11) > 0.3
    assert abs(np.median(c12[big] / c11[big]) - 0.8) < 0.02

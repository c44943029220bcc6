"""GPU tests of the p-MG base level's algebraic multigrid: the CSR kernels
against scipy, the assembled operator against the matrix-free one, the AMG
V-cycle (symmetry, convergence as a stationary iteration and as a PCG
preconditioner), and p-MG solves with [MULTIGRID COARSE SOLVER] AMG."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
sp = pytest.importorskip("scipy.sparse")


def lp():
    import libparanumal_amd
    return libparanumal_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def box(n, N, flag=1, curve=0.05, lam=0.3, **extra):
    s = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": n, "BOX NY": n + 1, "BOX NZ": n - 1,
         "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "BOX BOUNDARY FLAG": flag, "POLYNOMIAL DEGREE": N,
         "LAMBDA": lam, "DATA FILE": "data/ellipticSine3D.h", "BOX CURVED MAP AMPLITUDE": curve}
    s.update(extra)
    return s


def sync(el):
    import ctypes
    torch.cuda.synchronize()
    ctypes.CDLL("libamdhip64.so").hipStreamSynchronize(ctypes.c_void_p(el.stream))


@pytest.mark.parametrize("lanes", [1, 8, 64])
def test_csr_spmv_vs_scipy(lanes):
    rng = np.random.default_rng(lanes)
    A = sp.random(700, 500, density=0.05, format="csr", random_state=lanes, dtype=np.float64)
    A.sort_indices()
    x = rng.standard_normal(500)
    z = rng.standard_normal(700)
    rs, ci, v = dev(A.indptr.astype(np.int32)), dev(A.indices.astype(np.int32)), dev(A.data)
    xd, zd = dev(x), dev(z)
    y = torch.empty(700, dtype=torch.float64, device="cuda")
    L = lp().lib()
    s = torch.cuda.current_stream().cuda_stream
    lp().check(L.libp_csr_spmv_f64(700, rs.data_ptr(), ci.data_ptr(), v.data_ptr(), -1.5, xd.data_ptr(), 0.5,
                                   zd.data_ptr(), y.data_ptr(), lanes, s))
    torch.cuda.synchronize()
    ref = -1.5 * (A @ x) + 0.5 * z
    assert np.allclose(y.cpu().numpy(), ref, rtol=1e-13, atol=1e-13)
    # z aliasing y: y += A x
    lp().check(L.libp_csr_spmv_f64(700, rs.data_ptr(), ci.data_ptr(), v.data_ptr(), 1.0, xd.data_ptr(), 1.0,
                                   y.data_ptr(), y.data_ptr(), lanes, s))
    torch.cuda.synchronize()
    assert np.allclose(y.cpu().numpy(), ref + A @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("N,flag", [(1, 1), (2, 1), (3, 2)])
def test_assembled_operator_matches_matrix_free(N, flag):
    el = lp().Elliptic(box(3, N, flag))
    rs, ci, v = el.assembled_csr()
    n = el.geti("Ngather")
    A = sp.csr_matrix((v, ci, rs), shape=(n, n))
    print("csr", N, flag, n, A.nnz, flush=True)
    assert abs(A - A.T).max() <= 1e-13 * abs(A).max()
    x = np.random.default_rng(N).standard_normal(n)
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    el.operator(dev(x), y)
    sync(el)
    ref = A @ x
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.fixture(scope="module")
def amg12():
    el = lp().Elliptic(box(12, 1, 1, lam=0.0))
    el.amg_setup()
    rs, ci, v = el.assembled_csr()
    n = el.geti("Ngather")
    A = sp.csr_matrix((v, ci, rs), shape=(n, n))
    info = el.amg_info()
    print("amg levels", info, flush=True)
    return el, A, n, info


def vcycle(el, b):
    bd = dev(b)
    x = torch.empty_like(bd)
    el.amg_vcycle(bd, x)
    return x.cpu().numpy()


def test_amg_hierarchy(amg12):
    el, A, n, info = amg12
    assert info["levels"] >= 2
    assert info["rows"][0] == n
    assert all(a > b for a, b in zip(info["rows"], info["rows"][1:]))
    assert info["rows"][-1] <= 1000


def test_amg_vcycle_is_symmetric(amg12):
    """Chebyshev pre- and post-smoothing with one polynomial, R = P^T and an exact
    coarsest solve make the V-cycle a symmetric operator: x.V(y) = y.V(x)."""
    el, A, n, info = amg12
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    a, b = x @ vcycle(el, y), y @ vcycle(el, x)
    print("symmetry", a, b, flush=True)
    assert abs(a - b) <= 1e-10 * abs(a)


def test_amg_stationary_and_pcg_convergence(amg12):
    el, A, n, info = amg12
    rng = np.random.default_rng(4)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    norms = [np.linalg.norm(b)]
    for _ in range(8):
        x = x + vcycle(el, b - A @ x)
        norms.append(np.linalg.norm(b - A @ x))
    factors = [norms[k + 1] / norms[k] for k in range(len(norms) - 1)]
    print("stationary factors", factors, flush=True)
    assert max(factors[2:]) < 0.35
    # CG preconditioned by one V-cycle
    x = np.zeros(n); r = b.copy(); z = vcycle(el, r); p = z.copy(); rz = r @ z
    for it in range(1, 60):
        Ap = A @ p
        al = rz / (p @ Ap)
        x += al * p
        r -= al * Ap
        if np.linalg.norm(r) <= 1e-8 * norms[0]:
            break
        z = vcycle(el, r)
        rz, rzo = r @ z, rz
        p = z + (rz / rzo) * p
    print("AMG-PCG iterations", it, flush=True)
    assert it <= 15


def test_pmg_with_amg_base_level():
    """FPCG + p-MG 7,3,1 with the AMG base level reaches 1e-8 and the solution of
    the Jacobi-PCG base level."""
    base = box(4, 7, 1, lam=1.0, **{"LINEAR SOLVER": "FPCG", "PRECONDITIONER": "MULTIGRID",
                                    "MULTIGRID COARSENING": "CUSTOM", "MULTIGRID DEGREES": "7,3,1",
                                    "MULTIGRID PRECISION": "FP32", "LINEAR SOLVER STOPPING CRITERION": "RELATIVE",
                                    "MULTIGRID COARSE TOLERANCE": 1e-2})
    a = lp().Elliptic(dict(base, **{"MULTIGRID COARSE SOLVER": "AMG"}))
    b = lp().Elliptic(base)
    ra, rb = a.run(), b.run()
    print("amg base", ra, "jacobi-pcg base", rb, flush=True)
    assert ra["r_final"] <= 1e-8 * ra["r0"] * (1 + 1e-12)
    assert abs(ra["norm"] - rb["norm"]) < 1e-7 * rb["norm"]
    assert ra["iterations"] <= rb["iterations"] + 3


# ---------------------------------------------------------------------------
# against the oracle's restatement of parAlmond (oracle/amg.cpp, pinned in
# tests/test_oracle_amg.py): aggregates bit for bit, matrices and rho to rounding
# ---------------------------------------------------------------------------
def cube(n, **extra):
    s = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": n, "BOX NY": n, "BOX NZ": n,
         "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "BOX BOUNDARY FLAG": 1, "POLYNOMIAL DEGREE": 1,
         "LAMBDA": 0.0, "DATA FILE": "data/ellipticSine3D.h", "BOX CURVED MAP AMPLITUDE": 0.05}
    s.update(extra)
    return s


def compare_hierarchies(el, o, which):
    """Level by level: sizes, aggregates (bit for bit), A, P, R (1e-12) and rho (1e-10)."""
    nl = o.amg_levels(which)
    info = el.amg_info() if which == 0 else None
    if info is not None:
        assert info["levels"] == nl
    for l in range(nl):
        for m in ("A",) + (("P", "R") if l + 1 < nl else ()):
            rs, ci, v = (el.amg_level_array(l, m + "_" + k, which) for k in ("rs", "ci", "v"))
            ors, oci, ov = (o.amg_level_array(l, m + "_" + k, which) for k in ("rs", "ci", "v"))
            assert np.array_equal(rs, ors) and np.array_equal(ci, oci), (l, m)
            assert np.max(np.abs(v - ov)) <= 1e-12 * np.max(np.abs(ov)), (l, m)
        rho = el.amg_level_array(l, "rho", which)[0]
        assert abs(rho - o.amg_level_d(l, "rho", which)) <= 1e-10 * rho
        if l + 1 < nl:
            agg = el.amg_level_array(l, "aggregate", which)
            assert np.array_equal(agg, o.amg_level_array(l, "FineToCoarse", which).astype(np.int32)), l


@pytest.mark.parametrize("opts", [{}, {"PARALMOND AGGREGATION": "UNSMOOTHED", "PARALMOND SMOOTHER": "DAMPEDJACOBI"},
                                  {"PARALMOND STRENGTH": "RUGESTUBEN", "PARALMOND CHEBYSHEV DEGREE": 3}])
def test_amg_setup_and_cycle_match_oracle(opts):
    from tests.oracle_lib import OracleElliptic
    s = cube(16, **opts)
    o = OracleElliptic(s)
    o.amg_setup(0)
    el = lp().Elliptic(s)
    el.amg_setup()
    compare_hierarchies(el, o, 0)
    n = el.geti("Ngather")
    b = np.random.default_rng(4).standard_normal(n)
    # one smoothing from zero and from a guess, then the whole V-cycle
    for x0, zero in ((np.zeros(n), True), (np.random.default_rng(5).standard_normal(n), False)):
        xd = dev(x0)
        el.amg_smooth(0, dev(b), xd, zero)
        ref = o.amg_smooth(0, b, x0, zero)
        assert np.max(np.abs(xd.cpu().numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    el.amg_vcycle(dev(b), x)
    sync(el)
    ref = o.amg_vcycle(b)
    assert np.max(np.abs(x.cpu().numpy() - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_amg_rejects_unsupported_cycle():
    el = lp().Elliptic(cube(4, **{"PARALMOND CYCLE": "KCYCLE"}))
    with pytest.raises(lp().LibpError, match="VCYCLE only"):
        el.amg_setup()


def test_pmg_amg_base_matches_oracle_harness_settings():
    """testEllipticHex_C0_Multigrid as the reference composes it: the harness settings
    (test/testElliptic.py:34-72, PARALMOND AGGREGATION UNSMOOTHED) with the p-MG
    ending in parAlmond AMG. The base level's hierarchy matches the oracle's (the
    drand48 stream continues from the p-MG levels' Arnoldi starts), the solve takes
    the oracle's iterations +-1 and gives its norm to 1e-10."""
    from tests import oracle_lib
    from tests.oracle_lib import OracleElliptic
    from tests.test_oracle import ref_settings
    s = ref_settings(PRECONDITIONER="MULTIGRID", **{"LINEAR SOLVER": "FPCG", "MULTIGRID COARSE SOLVER": "AMG",
                                                    "MULTIGRID PRECISION": "FP64"})
    o = OracleElliptic(dict(s, **{"MULTIGRID COARSENING": "CUSTOM", "MULTIGRID DEGREES": "4,2,1"}))
    o.mg_setup()
    el = lp().Elliptic(dict(s, **{"MULTIGRID COARSENING": "CUSTOM", "MULTIGRID DEGREES": "4,2,1"}))
    compare_hierarchies(el, o, 1)
    onorm, oit, _ = oracle_lib.run(s)
    r = lp().Elliptic(s).run()
    print(f"harness p-MG + AMG: GPU {r['iterations']} its norm {r['norm']:.15f}; oracle {oit} its {onorm:.15f}",
          flush=True)
    assert abs(r["iterations"] - oit) <= 1
    assert abs(r["norm"] - onorm) <= 1e-10 * onorm
    assert abs(r["norm"] - 0.353553400508458) < 1e-5


def test_pmg_amg_base_hierarchy_and_vcycle_match_oracle():
    """Curved 12^3 box, p-MG 3 -> 1 in FP64 with the AMG base level (1331 rows, one
    coarsening): the base hierarchy matches the oracle's and so does one V-cycle."""
    from tests.oracle_lib import OracleElliptic
    s = cube(12, **{"POLYNOMIAL DEGREE": 3, "LAMBDA": 1.0, "PRECONDITIONER": "MULTIGRID", "LINEAR SOLVER": "FPCG",
                    "MULTIGRID COARSENING": "CUSTOM", "MULTIGRID DEGREES": "3,1", "MULTIGRID PRECISION": "FP64",
                    "MULTIGRID COARSE SOLVER": "AMG", "PARALMOND AGGREGATION": "UNSMOOTHED"})
    o = OracleElliptic(s)
    o.mg_setup()
    assert o.amg_levels(1) == 2
    el = lp().Elliptic(s)
    compare_hierarchies(el, o, 1)
    r = np.random.default_rng(8).standard_normal(o.geti("Ngather"))
    z = torch.zeros(r.size, dtype=torch.float64, device="cuda")
    el.precon(dev(r), z)
    sync(el)
    ref = o.mg_apply(r)
    assert np.max(np.abs(z.cpu().numpy() - ref)) <= 1e-10 * np.max(np.abs(ref))

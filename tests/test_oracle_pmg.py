"""CPU pins of the oracle's p-multigrid pieces against closed forms on small
dense systems (no GPU needed).

The oracle restates MGLevel (solvers/elliptic/src/ellipticPreconMultiGridLevel.cpp)
and the V-cycle (libs/parAlmond/parAlmondVcycle.cpp:34-60). Each test below
checks one restated function against mathematics that does not use the
restatement, chosen so that a plausible slip fails it:

- coarsen (:41-82) against the transpose of prolongate (:84-125): the
  inverse-multiplicity weighting makes R = P^T exactly; a dropped weight, a
  transposed 1-D matrix or a wrong sweep index breaks the equality;
- prolongate against polynomial exactness: the coarse interpolant of a
  polynomial of degree <= Nc per direction must come out as the fine
  interpolant (affine box);
- smoothChebyshev (:154-206) against the textbook Chebyshev error propagator
  E = T_{K+1}((theta - D^-1 A)/delta) / T_{K+1}(theta/delta) on [rho/10, rho]
  (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1);
- maxEigSmoothAx (:393-473) against Ritz values computed independently from
  the Krylov space of D^-1 A and the libc drand48 start vector, and against the
  dense spectrum;
- the V-cycle against the two-grid formula built from the dense pieces, and
  its symmetry x.V(y) = y.V(x) with an exact coarse solve.
"""
import ctypes

import numpy as np
import pytest

from tests.oracle_lib import OracleElliptic

O = OracleElliptic


def mg_settings(nx=3, N=4, degrees="4,2,1", curve=0.05, flag=1, lam=1, cheb=2, **kw):
    s = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": nx, "BOX NY": nx, "BOX NZ": nx,
         "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "BOX BOUNDARY FLAG": flag, "POLYNOMIAL DEGREE": N,
         "BOX CURVED MAP AMPLITUDE": curve, "LAMBDA": lam, "DATA FILE": "data/ellipticSine3D.h",
         "LINEAR SOLVER": "FPCG", "PRECONDITIONER": "MULTIGRID", "MULTIGRID COARSENING": "CUSTOM",
         "MULTIGRID DEGREES": degrees, "MULTIGRID SMOOTHER": "CHEBYSHEV", "MULTIGRID CHEBYSHEV DEGREE": cheb,
         "MULTIGRID COARSE TOLERANCE": 1e-14, "MULTIGRID COARSE MAXIT": 1000}
    s.update(kw)
    return s


def dense(apply, n_in):
    cols = []
    for j in range(n_in):
        e = np.zeros(n_in)
        e[j] = 1.0
        cols.append(apply(e))
    return np.array(cols).T


def setup_mg(**kw):
    o = OracleElliptic(mg_settings(**kw))
    o.mg_setup()
    return o


@pytest.fixture(scope="module")
def mg3():
    """3^3 curved box, N=4, p-MG 4 -> 2 -> 1, lambda = 1, Dirichlet."""
    return setup_mg()


def level_dense(o, l):
    n = o.mg_level_ndofs(l)
    A = dense(lambda e: o.mg_level_op(l, O.MG_OPERATOR, e), n)
    P = dense(lambda e: o.mg_level_op(l, O.MG_PROLONGATE, e, out=np.zeros(n)), o.mg_level_ndofs(l + 1))
    R = dense(lambda e: o.mg_level_op(l, O.MG_COARSEN, e), n)
    return A, P, R


@pytest.mark.parametrize("level", [0, 1])
def test_restriction_is_transpose_of_prolongation(mg3, level):
    A, P, R = level_dense(mg3, level)
    assert P.shape == (mg3.mg_level_ndofs(level), mg3.mg_level_ndofs(level + 1))
    assert np.max(np.abs(R - P.T)) <= 1e-14 * np.max(np.abs(P))
    # the weighting matters: without the inverse multiplicity (1/2, 1/4, 1/8 at
    # shared nodes) the restriction would be R W^-1, which is far from P^T
    mult = 1.0 / OracleElliptic(mg_settings(N=[4, 2][level], degrees=["4,1", "2,1"][level])).array("weightG")
    assert mult.max() == 8.0
    assert np.max(np.abs(R * mult[None, :] - P.T)) > 0.1 * np.max(np.abs(P))


def gathered_coords(o):
    """x, y, z at the gathered nodes of an oracle handle (first copy of each row)."""
    g2l = o.array("GlobalToLocal")
    out = []
    for c in ("x", "y", "z"):
        v = o.array(c)
        g = np.zeros(o.geti("Ngather"))
        g[g2l[g2l >= 0]] = v[g2l >= 0]
        out.append(g)
    return out


@pytest.mark.parametrize("degrees,fpoly", [
    ("4,2,1", lambda x, y, z: x * x * y - 0.5 * y * z * z + x * z + 0.3 + 0.7 * x * x * y * y * z * z),
    ("3,1", lambda x, y, z: 1 + x - 2 * y + 0.5 * x * y * z + 0.25 * y * z),
])
def test_prolongation_exact_on_coarse_polynomials(degrees, fpoly):
    """Affine box, Neumann faces, lambda = 1 (no masked node): prolongating the
    coarse interpolant of f in Q_Nc gives f at the fine nodes."""
    degs = [int(d) for d in degrees.split(",")]
    o = setup_mg(nx=2, N=degs[0], degrees=degrees, curve=0.0, flag=2)
    fine = gathered_coords(o)
    coarse_handle = OracleElliptic(mg_settings(nx=2, N=degs[1], degrees=f"{degs[1]},1", curve=0.0, flag=2))
    coarse = gathered_coords(coarse_handle)
    assert coarse_handle.geti("Ngather") == o.mg_level_ndofs(1)
    fc = fpoly(*coarse)
    Pf = o.mg_level_op(0, O.MG_PROLONGATE, fc, out=np.zeros(o.mg_level_ndofs(0)))
    ff = fpoly(*fine)
    assert np.max(np.abs(Pf - ff)) <= 1e-13 * np.max(np.abs(ff))
    # and a polynomial of one degree more in x is not reproduced (the check has teeth)
    g = lambda x, y, z: fpoly(x, y, z) + x ** (degs[1] + 1)
    Pg = o.mg_level_op(0, O.MG_PROLONGATE, g(*coarse), out=np.zeros(o.mg_level_ndofs(0)))
    assert np.max(np.abs(Pg - g(*fine))) > 1e-3


def chebyshev_propagator(M, lam0, lam1, K):
    """Textbook error propagator of K+1 Chebyshev steps on [lam0, lam1]."""
    theta, delta = 0.5 * (lam1 + lam0), 0.5 * (lam1 - lam0)
    n = M.shape[0]
    Y = (theta * np.eye(n) - M) / delta
    T0, T1 = np.eye(n), Y
    t0, t1 = 1.0, theta / delta
    for _ in range(K):
        T0, T1 = T1, 2 * Y @ T1 - T0
        t0, t1 = t1, 2 * (theta / delta) * t1 - t0
    return T1 / t1


@pytest.mark.parametrize("cheb", [1, 2, 3])
def test_chebyshev_smoother_is_textbook_polynomial(cheb):
    o = mg3_cheb(cheb)
    A, _, _ = level_dense(o, 0)
    n = A.shape[0]
    invD = o.mg_level_op(0, O.MG_INVDIAG, None)
    lam1, lam0 = o.mg_level_d(0, "lambda1"), o.mg_level_d(0, "lambda0")
    assert abs(lam0 - lam1 / 10) <= 1e-15 * lam1
    M = invD[:, None] * A
    # x = S r from a zero guess: error propagator I - S A
    S = dense(lambda e: o.mg_level_op(0, O.MG_SMOOTH, e, out=np.zeros(n), flag=1), n)
    E = chebyshev_propagator(M, lam0, lam1, cheb)
    assert np.max(np.abs(np.eye(n) - S @ A - E)) < 1e-11
    # from a nonzero guess: e1 = E e0
    rng = np.random.default_rng(3)
    r, x0 = rng.standard_normal(n), rng.standard_normal(n)
    xs = np.linalg.solve(A, r)
    x1 = o.mg_level_op(0, O.MG_SMOOTH, r, out=x0.copy(), flag=0)
    assert np.max(np.abs((x1 - xs) - E @ (x0 - xs))) < 1e-11 * np.max(np.abs(x0 - xs))
    # one polynomial degree less is far off (the comparison can fail)
    assert np.max(np.abs(np.eye(n) - S @ A - chebyshev_propagator(M, lam0, lam1, cheb - 1))) > 1e-2
    # on the spectrum of D^-1 A inside [lam0, lam1] the smoother contracts by the Chebyshev bound
    ev = np.linalg.eigvals(M).real
    assert ev.max() <= lam1 * (1 + 0.2)
    bound = 1.0 / np.cosh((cheb + 1) * np.arccosh((lam1 + lam0) / (lam1 - lam0)))
    inside = (ev >= lam0) & (ev <= lam1)
    Ev = np.abs(np.linalg.eigvals(E))
    assert inside.sum() > n // 2 and np.sort(Ev)[: inside.sum()].max() <= bound * (1 + 1e-8)


_cheb_cache = {}


def mg3_cheb(cheb):
    if cheb not in _cheb_cache:
        _cheb_cache[cheb] = setup_mg(nx=2, N=4, degrees="4,1", cheb=cheb)
    return _cheb_cache[cheb]


def libc_drand48(seed, n):
    libc = ctypes.CDLL(None)
    libc.drand48.restype = ctypes.c_double
    libc.srand48(ctypes.c_long(seed))
    return np.array([libc.drand48() for _ in range(n)])


@pytest.mark.parametrize("nx,N,curve", [(2, 4, 0.05), (3, 3, 0.0), (2, 2, 0.05)])
def test_max_eig_is_ritz_value_of_krylov_space(nx, N, curve):
    """rho from the 10-step Arnoldi equals the largest Ritz value of D^-1 A on
    span{v0, M v0, ..., M^9 v0} (v0 from libc's drand48 after srand48(0)),
    computed here by a QR of the Krylov matrix, and lies within 5% of the
    spectral radius."""
    o = OracleElliptic(mg_settings(nx=nx, N=N, curve=curve, degrees=f"{N},1"))
    n = o.geti("Ngather")
    A = dense(o.operator, n)
    M = A / o.diagonal()[:, None]
    k = min(10, n)
    v = libc_drand48(0, n)
    K = np.empty((n, k))
    K[:, 0] = v / np.linalg.norm(v)
    for j in range(1, k):
        w = M @ K[:, j - 1]
        K[:, j] = w / np.linalg.norm(w)
    Q, _ = np.linalg.qr(K)
    ritz = np.linalg.eigvals(Q.T @ M @ Q)
    rho_indep = np.max(np.abs(ritz))
    rho = o.max_eig(0)
    assert abs(rho - rho_indep) <= 1e-8 * rho_indep, (rho, rho_indep)
    lam = np.max(np.abs(np.linalg.eigvals(M)))
    assert 0.95 * lam <= rho <= lam * (1 + 1e-10)
    if n > 100:
        # Ritz values of a 9-dimensional space differ: the step count is pinned too
        # (the 27-dof box has too few distinct eigenvalues for that)
        Q9, _ = np.linalg.qr(K[:, :k - 1])
        assert abs(np.max(np.abs(np.linalg.eigvals(Q9.T @ M @ Q9))) - rho) > 1e-8 * rho


def test_mg_levels_use_the_arnoldi_rho():
    """The level's Chebyshev interval is [rho/10, rho] with rho from the drand48
    stream that srand48(0) starts (ellipticPreconMultiGridLevel.cpp:361-368)."""
    o = setup_mg(nx=2, N=4, degrees="4,1")
    probe = OracleElliptic(mg_settings(nx=2, N=4, degrees="4,1"))
    assert abs(o.mg_level_d(0, "lambda1") - probe.max_eig(0)) <= 1e-14 * probe.max_eig(0)


def test_two_grid_vcycle_matches_formula():
    """With degrees N,1 the V-cycle is B = S + P Ac^-1 R (I - A S) + S (I - A (S + P Ac^-1 R (I - A S)))
    (pre-smooth from zero, coarse correction, post-smooth), built from the dense
    level pieces and the coarse operator."""
    o = setup_mg(nx=3, N=3, degrees="3,1")
    A, P, R = level_dense(o, 0)
    n = A.shape[0]
    coarse = OracleElliptic(mg_settings(nx=3, N=1, degrees="1,1"))
    Ac = dense(coarse.operator, coarse.geti("Ngather"))
    S = dense(lambda e: o.mg_level_op(0, O.MG_SMOOTH, e, out=np.zeros(n), flag=1), n)
    C = P @ np.linalg.solve(Ac, R)
    I = np.eye(n)
    B = S + C @ (I - A @ S) + S @ (I - A @ (S + C @ (I - A @ S)))
    V = dense(o.mg_apply, n)
    assert np.max(np.abs(V - B)) <= 1e-9 * np.max(np.abs(B))
    # without the post-smoothing the formula is far off (the check has teeth)
    assert np.max(np.abs(V - (S + C @ (I - A @ S)))) > 1e-3 * np.max(np.abs(B))


def test_three_level_vcycle_symmetric(mg3):
    """x.V(y) = y.V(x) for the FP64 V-cycle with an exact coarse solve: the
    pre- and post-smoothers are the same symmetric polynomial in D^-1 A and
    R = P^T on every level."""
    n = mg3.mg_level_ndofs(0)
    rng = np.random.default_rng(7)
    for _ in range(3):
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        a, b = x @ mg3.mg_apply(y), y @ mg3.mg_apply(x)
        assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))
    # and positive definite
    assert x @ mg3.mg_apply(x) > 0

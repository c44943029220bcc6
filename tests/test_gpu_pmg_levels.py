"""GPU p-multigrid level operations, one at a time, against the CPU oracle.

The oracle's operations are pinned by closed forms in tests/test_oracle_pmg.py
(R = P^T, polynomial exactness of P, the textbook Chebyshev propagator, Ritz
values for lambda max, the two-grid formula). Here the HIP path of each one,
called through the C ABI (libp_elliptic_mg_level_*), must agree with the oracle
on the same inputs: FP64 to rounding, FP32 to single precision."""
import ctypes

import numpy as np
import pytest

from tests.oracle_lib import OracleElliptic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

O = OracleElliptic


def settings(prec, degrees="4,2,1", N=4, nx=3, **kw):
    s = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": nx, "BOX NY": nx + 1,
         "BOX NZ": nx, "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "BOX BOUNDARY FLAG": 1,
         "POLYNOMIAL DEGREE": N, "BOX CURVED MAP AMPLITUDE": 0.05, "LAMBDA": 1.0,
         "DATA FILE": "data/ellipticSine3D.h", "LINEAR SOLVER": "FPCG", "PRECONDITIONER": "MULTIGRID",
         "MULTIGRID COARSENING": "CUSTOM", "MULTIGRID DEGREES": degrees, "MULTIGRID PRECISION": prec,
         "MULTIGRID SMOOTHER": "CHEBYSHEV", "MULTIGRID CHEBYSHEV DEGREE": 2,
         "MULTIGRID COARSE TOLERANCE": 1e-14, "MULTIGRID COARSE MAXIT": 1000}
    s.update(kw)
    return s


def sync(p):
    ctypes.CDLL("libamdhip64.so").hipStreamSynchronize(ctypes.c_void_p(p.stream))


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


_cache = {}


def pair(prec, **kw):
    key = (prec, tuple(sorted(kw.items())))
    if key not in _cache:
        import libparanumal_amd as lp
        s = settings(prec, **kw)
        o = OracleElliptic(s)
        o.mg_setup()
        _cache[key] = (lp.Elliptic(s), o)
    return _cache[key]


def gpu_op(p, level, op, vin, n_out, out0=None, dtype=torch.float64, flag=0):
    vin_d = torch.from_numpy(np.ascontiguousarray(vin)).to(dtype).cuda() if vin is not None else None
    out = (torch.from_numpy(np.ascontiguousarray(out0)).to(dtype).cuda() if out0 is not None
           else torch.full((n_out,), float("nan"), dtype=dtype, device="cuda"))
    p.mg_level_op(level, op, vin_d, out, flag)
    sync(p)
    return out.double().cpu().numpy()


@pytest.mark.parametrize("prec,tol", [("FP64", 1e-12), ("FP32", 2e-5)])
@pytest.mark.parametrize("level", [0, 1])
def test_level_operations_match_oracle(prec, tol, level):
    p, o = pair(prec)
    dt = torch.float64 if prec == "FP64" else torch.float32
    n, bits = p.mg_level_ndofs(level)
    nc, _ = p.mg_level_ndofs(level + 1)
    assert bits == (64 if prec == "FP64" else 32)
    assert n == o.mg_level_ndofs(level) and nc == o.mg_level_ndofs(level + 1)
    rng = np.random.default_rng(100 + level)
    x, xc, y0 = rng.standard_normal(n), rng.standard_normal(nc), rng.standard_normal(n)
    if prec == "FP32":  # compare on inputs both sides see exactly
        x, xc, y0 = (v.astype(np.float32).astype(np.float64) for v in (x, xc, y0))
    # Chebyshev interval: rho from the same drand48 start, reductions in another order
    l1, l0 = p.mg_level_lambda(level)
    assert abs(l1 - o.mg_level_d(level, "lambda1")) <= 1e-10 * l1
    assert abs(l0 - l1 / 10) <= 1e-15 * l1
    # inverse diagonal
    assert relerr(gpu_op(p, level, O.MG_INVDIAG, None, n, dtype=dt), o.mg_level_op(level, O.MG_INVDIAG, None)) < tol
    # operator
    assert relerr(gpu_op(p, level, O.MG_OPERATOR, x, n, dtype=dt), o.mg_level_op(level, O.MG_OPERATOR, x)) < tol
    # coarsen (weighted restriction + coarse gather)
    assert relerr(gpu_op(p, level, O.MG_COARSEN, x, nc, dtype=dt), o.mg_level_op(level, O.MG_COARSEN, x)) < tol
    # prolongate: y0 += P xc
    ref = o.mg_level_op(level, O.MG_PROLONGATE, xc, out=y0.copy())
    assert relerr(gpu_op(p, level, O.MG_PROLONGATE, xc, n, out0=y0, dtype=dt), ref) < tol
    # Chebyshev smoother from zero and from y0
    ref = o.mg_level_op(level, O.MG_SMOOTH, x, out=np.zeros(n), flag=1)
    got = gpu_op(p, level, O.MG_SMOOTH, x, n, out0=np.zeros(n), dtype=dt, flag=1)
    assert relerr(got, ref) < 10 * tol
    ref = o.mg_level_op(level, O.MG_SMOOTH, x, out=y0.copy(), flag=0)
    got = gpu_op(p, level, O.MG_SMOOTH, x, n, out0=y0, dtype=dt, flag=0)
    assert relerr(got, ref) < 10 * tol


def test_restriction_is_transpose_of_prolongation_on_gpu():
    """R = P^T for the GPU transfers, as dense matrices on the 3x4x3 box 2 -> 1 level."""
    p, _ = pair("FP64")
    n, _ = p.mg_level_ndofs(1)
    nc, _ = p.mg_level_ndofs(2)
    P = np.array([gpu_op(p, 1, O.MG_PROLONGATE, np.eye(nc)[j], n, out0=np.zeros(n)) for j in range(nc)]).T
    R = np.array([gpu_op(p, 1, O.MG_COARSEN, np.eye(n)[j], nc) for j in range(n)]).T
    assert np.max(np.abs(R - P.T)) <= 1e-14 * np.max(np.abs(P))


def test_fp64_vcycle_symmetric_on_gpu():
    """x.V(y) = y.V(x) for the GPU FP64 V-cycle with an exact coarse solve."""
    p, o = pair("FP64")
    n = o.geti("Ngather")
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal(n), rng.standard_normal(n)

    def V(r):
        rd = torch.from_numpy(r).cuda()
        z = torch.zeros_like(rd)
        p.precon(rd, z)
        sync(p)
        return z.cpu().numpy()

    a, b = x @ V(y), y @ V(x)
    assert abs(a - b) <= 1e-11 * max(abs(a), abs(b))
    assert relerr(V(y), o.mg_apply(y)) < 1e-10

"""The benchmark configurations' solves on the GPU, against the CPU oracle.

Configuration 2 (affine hex box, N=7, lambda=0, PCG+Jacobi) and configuration 4
(curved hex box, N=15, lambda=1, PCG+Jacobi on the FP64 MFMA operator) run
through elliptic_t::Run (solvers/elliptic/src/ellipticRun.cpp:30-248) and
LinearSolver::pcg (libs/linearSolver/linearSolverPCG.cpp:35-171). At sizes the
oracle finishes in seconds, iteration counts must agree within one and the
solution norms to 1e-10; at configuration 2's full size (48^3 elements,
37,595,375 gathered dofs) the checks are size-independent: the dof count,
operator symmetry, both stopping rules and the error against the manufactured
solution."""
import ctypes

import numpy as np
import pytest

from tests import oracle_lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def box(nx, N, curve=0.0, lam=0.0, ny=None, nz=None, **kw):
    s = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": nx, "BOX NY": ny or nx,
         "BOX NZ": nz or nx, "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "BOX BOUNDARY FLAG": 1,
         "POLYNOMIAL DEGREE": N, "BOX CURVED MAP AMPLITUDE": curve, "LAMBDA": lam,
         "DATA FILE": "data/ellipticSine3D.h", "DISCRETIZATION": "CONTINUOUS", "LINEAR SOLVER": "PCG",
         "PRECONDITIONER": "JACOBI"}
    s.update(kw)
    return s


def sync(p):
    ctypes.CDLL("libamdhip64.so").hipStreamSynchronize(ctypes.c_void_p(p.stream))


def solve_vs_oracle(s):
    import libparanumal_amd as lp
    onorm, oit, ng = oracle_lib.run(s)
    el = lp.Elliptic(s)
    assert el.geti("Ngather") == ng
    r = el.run()
    print(f"{s['BOX NX']}x{s['BOX NY']}x{s['BOX NZ']} N={s['POLYNOMIAL DEGREE']}: GPU {r['iterations']} its "
          f"norm {r['norm']:.15f}, oracle {oit} its norm {onorm:.15f}, error {r['error_l2']:.3e}", flush=True)
    assert abs(r["iterations"] - oit) <= 1
    assert abs(r["norm"] - onorm) <= 1e-10 * abs(onorm)
    el.close()
    return r, oit


@pytest.mark.parametrize("stop", ["ABS/REL-INITRESID", "RELATIVE"])
def test_c2_pcg_jacobi_n7_vs_oracle(stop):
    """Configuration 2's solve (affine box, N=7, lambda=0, PCG+Jacobi) at 10^3 elements
    (328,509 gathered dofs) under both stopping rules."""
    s = box(10, 7, **{"LINEAR SOLVER STOPPING CRITERION": stop})
    r, _ = solve_vs_oracle(s)
    if stop == "RELATIVE":
        assert r["r_final"] <= 1e-8 * r["r0"] * (1 + 1e-12)
    else:
        assert r["r_final"] <= max(1e-8 * r["r0"], 1e-8) * (1 + 1e-12)
    assert r["error_l2"] < 1e-8


def test_c4_pcg_jacobi_n15_curved_vs_oracle():
    """Configuration 4's solve (curved map, N=15, lambda=1, PCG+Jacobi on the MFMA
    operator) on a 3x2x2 box (37,004 gathered dofs)."""
    s = box(3, 15, curve=0.05, lam=1.0, ny=2, nz=2)
    r, _ = solve_vs_oracle(s)
    assert r["r_final"] <= max(1e-8 * r["r0"], 1e-8) * (1 + 1e-12)


def test_c2_full_size_properties():
    """48^3 elements, N=7: the gathered dof count (48*7-1)^3, x.Ay = y.Ax for
    the full operator, and the PCG+Jacobi solve under both stopping rules with the
    error against the manufactured solution."""
    import libparanumal_amd as lp
    s = box(48, 7)
    el = lp.Elliptic(s)
    Ng = el.geti("Ngather")
    assert Ng == (48 * 7 - 1) ** 3 == 37_595_375
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(Ng, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    y = torch.rand(Ng, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    Ax, Ay = torch.empty_like(x), torch.empty_like(y)
    el.operator(x, Ax)
    el.operator(y, Ay)
    sync(el)
    a, b = float(x @ Ay), float(y @ Ax)
    scale = float(torch.linalg.norm(x) * torch.linalg.norm(Ay))
    print(f"48^3 symmetry: x.Ay {a:.17e} y.Ax {b:.17e} |x||Ay| {scale:.3e}", flush=True)
    assert abs(a - b) <= 1e-12 * scale
    assert float(x @ Ax) > 0 and float(y @ Ay) > 0
    del x, y, Ax, Ay
    r = el.run()
    print(f"48^3 PCG+Jacobi (reference rule): {r['iterations']} its, r0 {r['r0']:.4e}, r {r['r_final']:.4e}, "
          f"error {r['error_l2']:.3e}, norm {r['norm']:.15f}", flush=True)
    assert r["r_final"] <= max(1e-8 * r["r0"], 1e-8) * (1 + 1e-12)
    assert r["error_l2"] < 1e-9
    assert abs(r["norm"] - 1.0) < 1e-9  # |u|_M of sin(pi x)sin(pi y)sin(pi z) on [-1,1]^3 is 1
    el.close()
    el = lp.Elliptic(dict(s, **{"LINEAR SOLVER STOPPING CRITERION": "RELATIVE"}))
    r2 = el.run()
    print(f"48^3 PCG+Jacobi (relative 1e-8): {r2['iterations']} its, r {r2['r_final']:.4e}, "
          f"error {r2['error_l2']:.3e}", flush=True)
    assert r2["r_final"] <= 1e-8 * r2["r0"] * (1 + 1e-12)
    assert r2["iterations"] > r["iterations"]
    assert r2["error_l2"] < 1e-9
    el.close()

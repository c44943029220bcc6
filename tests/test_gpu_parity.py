"""GPU parity tests: the HIP path (through the C ABI) against the CPU oracle."""
import numpy as np
import pytest

from tests import oracle_lib
from tests.oracle_lib import OracleElliptic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def L():
    import libparanumal_amd
    return libparanumal_amd.lib()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def stream():
    return torch.cuda.current_stream().cuda_stream


def box(n, N, flag=1, curve=0.0, lam=1.0, **extra):
    s = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": n, "BOX NY": n + 1, "BOX NZ": n - 1,
         "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "BOX BOUNDARY FLAG": flag, "POLYNOMIAL DEGREE": N,
         "LAMBDA": lam, "DATA FILE": "data/ellipticSine3D.h", "BOX CURVED MAP AMPLITUDE": curve}
    s.update(extra)
    return s


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ---------------------------------------------------------------------------
# Ax kernel in isolation: oracle arrays in, compare to the oracle's OKL restatement
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N", list(range(1, 16)))
def test_ax_kernel_f64(N):
    n = 3 if N <= 8 else 2
    o = OracleElliptic(box(n, N, curve=0.05, lam=0.7))
    rng = np.random.default_rng(N)
    q = rng.uniform(-1, 1, o.geti("Ngather"))
    ref = o.partial_ax(q)
    E = o.geti("Nelements")
    g2l, wJ, ggeo, D = dev(o.array("GlobalToLocal")), dev(o.array("wJ")), dev(o.array("ggeo")), dev(o.array("D"))
    qd = dev(q)
    out = torch.full((ref.size,), np.nan, dtype=torch.float64, device="cuda")
    st = L().libp_ax_hex_partial_f64(N, E, None, g2l.data_ptr(), wJ.data_ptr(), ggeo.data_ptr(), D.data_ptr(), 0.7,
                                     qd.data_ptr(), out.data_ptr(), stream())
    assert st == 0
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("N", [1, 3, 7, 11])
def test_ax_kernel_lambda_zero(N):
    """lambda = 0 (the Poisson operator) takes the kernel instance that does not read
    wJ: it matches the oracle to 1e-12, and a wJ full of NaN changes nothing (any read
    of it would turn the result into NaN). FP32 agrees with FP64 to FP32 rounding."""
    n = 3 if N <= 8 else 2
    o = OracleElliptic(box(n, N, curve=0.05, lam=0.0))
    rng = np.random.default_rng(100 + N)
    q = rng.uniform(-1, 1, o.geti("Ngather"))
    ref = o.partial_ax(q)
    E = o.geti("Nelements")
    g2l, ggeo, D = dev(o.array("GlobalToLocal")), dev(o.array("ggeo")), dev(o.array("D"))
    wJnan = torch.full((o.array("wJ").size,), np.nan, dtype=torch.float64, device="cuda")
    qd = dev(q)
    out = torch.full((ref.size,), 7.0, dtype=torch.float64, device="cuda")
    st = L().libp_ax_hex_partial_f64(N, E, None, g2l.data_ptr(), wJnan.data_ptr(), ggeo.data_ptr(), D.data_ptr(),
                                     0.0, qd.data_ptr(), out.data_ptr(), stream())
    assert st == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.isfinite(got).all()
    assert relerr(got, ref) < 1e-12
    out32 = torch.full((ref.size,), 7.0, dtype=torch.float32, device="cuda")
    wJ32, ggeo32, D32, q32 = wJnan.float(), ggeo.float(), D.float(), qd.float()  # held until the kernel has run
    st = L().libp_ax_hex_partial_f32(N, E, None, g2l.data_ptr(), wJ32.data_ptr(), ggeo32.data_ptr(), D32.data_ptr(),
                                     0.0, q32.data_ptr(), out32.data_ptr(), stream())
    assert st == 0
    torch.cuda.synchronize()
    got32 = out32.cpu().numpy().astype(np.float64)
    assert np.isfinite(got32).all()
    assert relerr(got32, ref) < 1e-5


@pytest.mark.parametrize("use_list", [False, True])
def test_ax_mfma_n15(use_list):
    o = OracleElliptic(box(2, 15, curve=0.05, lam=0.9))
    rng = np.random.default_rng(15)
    q = rng.uniform(-1, 1, o.geti("Ngather"))
    ref = o.partial_ax(q)
    E = o.geti("Nelements")
    g2l, wJ, ggeo, D = dev(o.array("GlobalToLocal")), dev(o.array("wJ")), dev(o.array("ggeo")), dev(o.array("D"))
    qd = dev(q)
    out = torch.full((ref.size,), np.nan, dtype=torch.float64, device="cuda")
    lst = dev(np.arange(E, dtype=np.int32)[::-1].copy()) if use_list else None
    st = L().libp_ax_hex_partial_mfma_n15(E, lst.data_ptr() if use_list else None, g2l.data_ptr(), wJ.data_ptr(),
                                          ggeo.data_ptr(), D.data_ptr(), 0.9, qd.data_ptr(), out.data_ptr(), stream())
    assert st == 0
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), ref) < 1e-12


def test_ax_kernel_element_list_and_tail():
    # odd element count (tail launch for NE>1) and an explicit element list
    N = 2
    o = OracleElliptic(box(4, N, lam=1.3))
    q = np.random.default_rng(1).uniform(-1, 1, o.geti("Ngather"))
    ref = o.partial_ax(q)
    E = o.geti("Nelements")
    Np = (N + 1) ** 3
    sub = np.arange(E - 3, dtype=np.int32)[::-1].copy()  # all but the last 3, reversed
    args = [dev(o.array(k)) for k in ("GlobalToLocal", "wJ", "ggeo", "D")]
    qd = dev(q)
    out = torch.zeros(ref.size, dtype=torch.float64, device="cuda")
    subd = dev(sub)
    st = L().libp_ax_hex_partial_f64(N, len(sub), subd.data_ptr(), *[a.data_ptr() for a in args], 1.3, qd.data_ptr(),
                                     out.data_ptr(), stream())
    assert st == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(E, Np)
    refe = ref.reshape(E, Np)
    assert relerr(got[sub], refe[sub]) < 1e-12
    assert np.all(got[E - 3:] == 0)


def test_zero_length_inputs():
    """Empty inputs through the C ABI: zero elements and zero rows write nothing
    and return OK, and a reduction over zero entries writes 0 (not a stale value)."""
    args = [torch.zeros(8, dtype=torch.float64, device="cuda") for _ in range(4)]
    g2l = torch.zeros(8, dtype=torch.int32, device="cuda")
    out = torch.full((8,), 7.0, dtype=torch.float64, device="cuda")
    assert L().libp_ax_hex_partial_f64(7, 0, None, g2l.data_ptr(), *[a.data_ptr() for a in args[:3]], 1.0,
                                       args[3].data_ptr(), out.data_ptr(), stream()) == 0
    rs = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert L().libp_ogs_gather_f64(0, rs.data_ptr(), g2l.data_ptr(), args[0].data_ptr(), out.data_ptr(),
                                   stream()) == 0
    ws = torch.zeros(L().libp_reduce_ws_bytes(), dtype=torch.uint8, device="cuda")
    res = torch.full((1,), 5.0, dtype=torch.float64, device="cuda")
    assert L().libp_inner_prod_f64(0, args[0].data_ptr(), args[1].data_ptr(), res.data_ptr(), ws.data_ptr(),
                                   stream()) == 0
    torch.cuda.synchronize()
    assert torch.all(out == 7.0)
    assert res.item() == 0.0

@pytest.mark.parametrize("N", list(range(1, 16)))
def test_ax_kernel_f32(N):
    """Every degree of the FP32 kernel (the p-MG levels run whichever degrees
    the coarsening picks) against the FP64 oracle, to FP32 rounding."""
    o = OracleElliptic(box(3 if N <= 8 else 2, N, curve=0.05))
    q = np.random.default_rng(2).uniform(-1, 1, o.geti("Ngather"))
    ref = o.partial_ax(q)
    E = o.geti("Nelements")
    f = lambda k: dev(o.array(k).astype(np.float32))
    g2l = dev(o.array("GlobalToLocal"))
    wJ, ggeo, D = f("wJ"), f("ggeo"), f("D")
    qd = dev(q.astype(np.float32))
    out = torch.zeros(ref.size, dtype=torch.float32, device="cuda")
    st = L().libp_ax_hex_partial_f32(N, E, None, g2l.data_ptr(), wJ.data_ptr(), ggeo.data_ptr(), D.data_ptr(), 1.0,
                                     qd.data_ptr(), out.data_ptr(), stream())
    assert st == 0
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy().astype(np.float64), ref) < 2e-5


def test_ax_bad_degree_rejected():
    st = L().libp_ax_hex_partial_f64(16, 1, None, None, None, None, None, 0.0, None, None, None)
    assert st != 0


# ---------------------------------------------------------------------------
# gather / scatter: bit-exact
# ---------------------------------------------------------------------------
def test_gather_bitexact():
    o = OracleElliptic(box(3, 4, curve=0.03))
    rng = np.random.default_rng(5)
    vL = rng.standard_normal(o.geti("Nelements") * o.geti("Np"))
    for trans, rsn, cin in ((True, "rowStartsT", "colIdsT"), (False, "rowStartsN", "colIdsN")):
        ref = o.gather(vL, trans)
        rs, ci = dev(o.array(rsn)), dev(o.array(cin))
        v = dev(vL)
        gv = torch.zeros(ref.size, dtype=torch.float64, device="cuda")
        assert L().libp_ogs_gather_f64(ref.size, rs.data_ptr(), ci.data_ptr(), v.data_ptr(), gv.data_ptr(),
                                       stream()) == 0
        torch.cuda.synchronize()
        assert np.array_equal(gv.cpu().numpy(), ref)


def test_fused_gather_pap_and_jacobi_update():
    """The fused PCG entry points through the C ABI. libp_pcg_gather_pap_f64:
    gv bit-identical to the oracle's gather, p.gv and alpha = rdotz1/p.gv
    against numpy. libp_pcg_jacobi_update_f64: x, r, z entry by entry as the
    formulas (within 2 ulps: the compiler may fuse a*b+c; z = invD*r exactly
    from the kernel's own r), r.r and r.z against numpy,
    ITERS counted, RDOTZ shifted and beta = r.z/old r.z. A planted STOP flag
    leaves every output untouched. Ragged sizes: 4x5x3 elements, N=4."""
    o = OracleElliptic(box(4, 4, curve=0.03))
    rng = np.random.default_rng(8)
    vL = rng.standard_normal(o.geti("Nelements") * o.geti("Np"))
    ref = o.gather(vL, True)
    n = ref.size
    p = rng.standard_normal(n)
    rs, ci, v, pd = dev(o.array("rowStartsT")), dev(o.array("colIdsT")), dev(vL), dev(p)
    ws = torch.zeros(L().libp_reduce_ws_bytes(), dtype=torch.uint8, device="cuda")
    s = np.zeros(10)
    s[0] = 2.5  # RDOTZ1
    sd = dev(s)
    gv = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert L().libp_pcg_gather_pap_f64(n, rs.data_ptr(), ci.data_ptr(), v.data_ptr(), gv.data_ptr(), pd.data_ptr(),
                                       sd.data_ptr(), ws.data_ptr(), stream()) == 0
    torch.cuda.synchronize()
    assert np.array_equal(gv.cpu().numpy(), ref)
    sh = sd.cpu().numpy()
    pap = float(p @ ref)
    assert abs(sh[4] - pap) <= 1e-13 * abs(p) @ np.abs(ref)
    assert abs(sh[2] - 2.5 / sh[4]) <= 1e-15 * abs(sh[2])
    # fused Jacobi update with alpha = sh[2]
    x, r, invD = rng.standard_normal(n), rng.standard_normal(n), rng.uniform(0.5, 2, n)
    xd, rd, zd, Ad, dd = dev(x), dev(r), torch.zeros(n, dtype=torch.float64, device="cuda"), gv, dev(invD)
    s2 = sh.copy()
    s2[7] = 1e-300  # TOL
    sd2 = dev(s2)
    assert L().libp_pcg_jacobi_update_f64(n, dd.data_ptr(), pd.data_ptr(), Ad.data_ptr(), xd.data_ptr(), rd.data_ptr(),
                                          zd.data_ptr(), sd2.data_ptr(), ws.data_ptr(), stream()) == 0
    torch.cuda.synchronize()
    alpha = s2[2]
    eps = np.finfo(np.float64).eps
    xg, rg, zg = xd.cpu().numpy(), rd.cpu().numpy(), zd.cpu().numpy()
    assert np.all(np.abs(xg - (x + alpha * p)) <= 2 * eps * (np.abs(x) + np.abs(alpha * p)))
    assert np.all(np.abs(rg - (r - alpha * ref)) <= 2 * eps * (np.abs(r) + np.abs(alpha * ref)))
    assert np.array_equal(zg, invD * rg)
    rr, zr = rg, zg
    out = sd2.cpu().numpy()
    assert abs(out[5] - rr @ rr) <= 1e-13 * (rr @ rr)
    assert abs(out[0] - rr @ zr) <= 1e-13 * abs(rr) @ np.abs(zr)
    assert out[1] == s2[0] and out[9] == 1.0 and out[8] == 0.0
    assert abs(out[3] - out[0] / out[1]) <= 1e-15 * abs(out[3])
    # a raised STOP flag: both entries return without touching anything
    out[8] = 1.0
    sd3 = dev(out)
    before = [t.clone() for t in (xd, rd, zd, gv)]
    L().libp_pcg_gather_pap_f64(n, rs.data_ptr(), ci.data_ptr(), v.data_ptr(), gv.data_ptr(), pd.data_ptr(),
                                sd3.data_ptr(), ws.data_ptr(), stream())
    L().libp_pcg_jacobi_update_f64(n, dd.data_ptr(), pd.data_ptr(), Ad.data_ptr(), xd.data_ptr(), rd.data_ptr(),
                                   zd.data_ptr(), sd3.data_ptr(), ws.data_ptr(), stream())
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(before, (xd, rd, zd, gv)))
    assert np.array_equal(sd3.cpu().numpy(), out)

@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fused_gather_coarsen_assemble_match_two_calls(prec):
    """The fused forms used inside the V-cycle are bit-identical to the two
    calls they replace: gather-residual (r - gather) and gather-add (y +=
    gather) against gather + axpy; the weighted coarsening against amxpy +
    coarsening (a random transfer matrix: both paths use the same one); the
    Chebyshev start and step in the gather against gather + cheb kernels; the
    accumulating learned assembly against assembly + axpy (FP32 only)."""
    dt = torch.float64 if prec == "f64" else torch.float32
    npdt = np.float64 if prec == "f64" else np.float32
    lib = L()
    fn = lambda name: getattr(lib, f"libp_{name}_{prec}")
    o = OracleElliptic(box(3, 4, curve=0.03))
    rng = np.random.default_rng(31)
    E, Np = o.geti("Nelements"), o.geti("Np")
    n = o.geti("Ngather")
    rs, ci = dev(o.array("rowStartsT")), dev(o.array("colIdsT"))
    vL = dev(rng.standard_normal(E * Np).astype(npdt))
    r = dev(rng.standard_normal(n).astype(npdt))
    y0 = rng.standard_normal(n).astype(npdt)
    st = stream()
    # residual
    a = torch.empty(n, dtype=dt, device="cuda")
    b = torch.empty(n, dtype=dt, device="cuda")
    assert fn("ogs_gather")(n, rs.data_ptr(), ci.data_ptr(), vL.data_ptr(), a.data_ptr(), st) == 0
    assert fn("axpy")(n, 1.0, r.data_ptr(), -1.0, a.data_ptr(), st) == 0
    assert fn("ogs_gather_residual")(n, rs.data_ptr(), ci.data_ptr(), vL.data_ptr(), r.data_ptr(), b.data_ptr(),
                                      st) == 0
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    # add
    a, b = dev(y0), dev(y0)
    g = torch.empty(n, dtype=dt, device="cuda")
    assert fn("ogs_gather")(n, rs.data_ptr(), ci.data_ptr(), vL.data_ptr(), g.data_ptr(), st) == 0
    assert fn("axpy")(n, 1.0, g.data_ptr(), 1.0, a.data_ptr(), st) == 0
    assert fn("ogs_gather_add")(n, rs.data_ptr(), ci.data_ptr(), vL.data_ptr(), b.data_ptr(), st) == 0
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    # weighted coarsening 4 -> 2
    Nf, Nc = 4, 2
    P = dev(rng.standard_normal((Nf + 1) * (Nc + 1)).astype(npdt))
    w = dev(rng.uniform(0.1, 1.0, n).astype(npdt))
    qf = dev(rng.standard_normal(n).astype(npdt))
    g2l = dev(o.array("GlobalToLocal"))
    wq = torch.empty(n, dtype=dt, device="cuda")
    c1 = torch.empty(E * (Nc + 1) ** 3, dtype=dt, device="cuda")
    c2 = torch.empty_like(c1)
    assert fn("amxpy")(n, 1.0, w.data_ptr(), qf.data_ptr(), 0.0, wq.data_ptr(), st) == 0
    assert fn("mg_coarsen_hex")(Nf, Nc, E, None, g2l.data_ptr(), P.data_ptr(), wq.data_ptr(), c1.data_ptr(), st) == 0
    assert fn("mg_coarsen_weighted_hex")(Nf, Nc, E, None, g2l.data_ptr(), P.data_ptr(), w.data_ptr(), qf.data_ptr(),
                                         c2.data_ptr(), st) == 0
    torch.cuda.synchronize()
    assert torch.equal(c1, c2)
    # Chebyshev start (with A x) and step fused into the gather
    iD = dev(rng.uniform(0.5, 2, n).astype(npdt))
    c1s, c2s = 0.7, -0.3
    for start in (True, False):
        R0, d0, X0 = (rng.standard_normal(n).astype(npdt) for _ in range(3))
        Ra, da, Xa = dev(R0), dev(d0), dev(X0)
        Rb, db, Xb = dev(R0), dev(d0), dev(X0)
        assert fn("ogs_gather")(n, rs.data_ptr(), ci.data_ptr(), vL.data_ptr(), g.data_ptr(), st) == 0
        if start:
            assert fn("cheb_start")(n, iD.data_ptr(), r.data_ptr(), g.data_ptr(), c1s, Ra.data_ptr(), da.data_ptr(),
                                    Xa.data_ptr(), 0, st) == 0
            assert fn("ogs_gather_cheb_start")(n, rs.data_ptr(), ci.data_ptr(), vL.data_ptr(), iD.data_ptr(),
                                               r.data_ptr(), c1s, Rb.data_ptr(), db.data_ptr(), Xb.data_ptr(), st) == 0
        else:
            assert fn("cheb_step")(n, iD.data_ptr(), g.data_ptr(), c1s, c2s, Ra.data_ptr(), da.data_ptr(),
                                   Xa.data_ptr(), st) == 0
            assert fn("ogs_gather_cheb_step")(n, rs.data_ptr(), ci.data_ptr(), vL.data_ptr(), iD.data_ptr(), c1s, c2s,
                                              Rb.data_ptr(), db.data_ptr(), Xb.data_ptr(), st) == 0
        torch.cuda.synchronize()
        assert torch.equal(Ra, Rb) and torch.equal(da, db) and torch.equal(Xa, Xb), start
    if prec == "f32":
        Y = dev(rng.standard_normal(E * Np).astype(np.float32))
        sc = dev(rng.uniform(0.5, 2, E).astype(np.float32))
        invD = dev(rng.uniform(0.5, 2, n).astype(np.float32))
        a, b = dev(y0), dev(y0)
        cc = torch.empty(n, dtype=dt, device="cuda")
        assert lib.libp_learned_assemble_f32(n, rs.data_ptr(), ci.data_ptr(), Np, Y.data_ptr(), Np, sc.data_ptr(),
                                             None, r.data_ptr(), invD.data_ptr(), cc.data_ptr(), st) == 0
        assert lib.libp_axpy_f32(n, 1.0, cc.data_ptr(), 1.0, a.data_ptr(), st) == 0
        assert lib.libp_learned_assemble_add_f32(n, rs.data_ptr(), ci.data_ptr(), Np, Y.data_ptr(), Np, sc.data_ptr(),
                                                 None, r.data_ptr(), invD.data_ptr(), b.data_ptr(), st) == 0
        # Y already scaled per element (what libp_mlp_linear_fwd_scaled writes), scale NULL: the same sums
        Ys = (Y.view(E, Np) * sc.view(E, 1)).contiguous()
        cs = torch.empty(n, dtype=dt, device="cuda")
        assert lib.libp_learned_assemble_f32(n, rs.data_ptr(), ci.data_ptr(), Np, Ys.data_ptr(), Np, None,
                                             None, r.data_ptr(), invD.data_ptr(), cs.data_ptr(), st) == 0
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        assert torch.equal(cs, cc)

def test_rdotz_from_f32_matches_convert_and_rdotz():
    """The FP32 preconditioner output widened in the pass that forms r.z and
    z.Ap (libp_pcg_rdotz_from_f32_f64) against libp_convert_f32_to_f64 followed
    by libp_pcg_rdotz_f64: z and all four scalars they write are bit-identical,
    flexible or not, at a ragged size."""
    lib = L()
    n = 123457
    rng = np.random.default_rng(41)
    r, Ap = dev(rng.standard_normal(n)), dev(rng.standard_normal(n))
    z32 = dev(rng.standard_normal(n).astype(np.float32))
    st = stream()
    for flexible in (0, 1):
        ws = torch.zeros(lib.libp_reduce_ws_bytes(), dtype=torch.uint8, device="cuda")
        s0 = np.zeros(10)
        s0[0], s0[2] = 1.7, 0.3
        sa, sb = dev(s0), dev(s0)
        za = torch.empty(n, dtype=torch.float64, device="cuda")
        zb = torch.empty_like(za)
        assert lib.libp_convert_f32_to_f64(n, z32.data_ptr(), za.data_ptr(), st) == 0
        assert lib.libp_pcg_rdotz_f64(n, r.data_ptr(), za.data_ptr(), Ap.data_ptr(), flexible, sa.data_ptr(),
                                      ws.data_ptr(), st) == 0
        assert lib.libp_pcg_rdotz_from_f32_f64(n, r.data_ptr(), z32.data_ptr(), zb.data_ptr(), Ap.data_ptr(), flexible,
                                               sb.data_ptr(), ws.data_ptr(), st) == 0
        torch.cuda.synchronize()
        assert torch.equal(za, zb)
        assert np.array_equal(sa.cpu().numpy(), sb.cpu().numpy()), flexible

def test_scatter_bitexact():
    o = OracleElliptic(box(3, 3))
    g = np.random.default_rng(6).standard_normal(o.geti("Ngather"))
    ref = o.scatter(g)
    rs, ci = dev(o.array("rowStartsT")), dev(o.array("colIdsT"))
    gd = dev(g)
    v = torch.zeros(ref.size, dtype=torch.float64, device="cuda")
    assert L().libp_ogs_scatter_f64(g.size, rs.data_ptr(), ci.data_ptr(), gd.data_ptr(), v.data_ptr(), stream()) == 0
    torch.cuda.synchronize()
    assert np.array_equal(v.cpu().numpy(), ref)


# ---------------------------------------------------------------------------
# product setup vs oracle: index maps bit-exact, geometry to rounding
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,N,flag,curve", [(3, 4, 1, 0.0), (4, 2, 1, 0.05), (4, 3, -1, 0.0), (2, 7, 1, 0.05),
                                            (3, 1, 1, 0.0)])
def test_setup_parity(n, N, flag, curve):
    import libparanumal_amd as lp
    s = box(n, N, flag, curve, PRECONDITIONER="JACOBI")
    o = OracleElliptic(s)
    p = lp.Elliptic(s)
    assert p.geti("Ngather") == o.geti("Ngather")
    for name in ("globalIds", "maskedGlobalIds", "GlobalToLocal", "rowStartsT", "colIdsT", "rowStartsN", "colIdsN",
                 "blockRowStartsT", "blockRowStartsN", "mapB"):
        a, b = p.host_array(name), o.array(name)
        assert a.dtype.itemsize == b.dtype.itemsize or name in ("mapB",)
        assert np.array_equal(a.astype(np.int64), b.astype(np.int64)), name
    for name, tol in (("x", 1e-14), ("y", 1e-14), ("z", 1e-14), ("wJ", 1e-13), ("ggeo", 1e-12), ("D", 1e-13),
                      ("gllz", 1e-15), ("gllw", 1e-14), ("weightG", 0.0)):
        a, b = p.host_array(name), o.array(name)
        assert relerr(a, b) <= tol, name
    assert relerr(1.0 / p.host_array("invDiag"), o.diagonal()) < 1e-12


def test_operator_parity():
    import libparanumal_amd as lp
    s = box(3, 5, 1, 0.05, lam=0.3)
    o = OracleElliptic(s)
    p = lp.Elliptic(s)
    q = np.random.default_rng(7).uniform(-1, 1, o.geti("Ngather"))
    ref = o.operator(q)
    qd = dev(q)
    Aq = torch.zeros(ref.size, dtype=torch.float64, device="cuda")
    p.operator(qd, Aq)
    torch.cuda.synchronize()
    lp.hip_d2h(0, 0, 0, p.stream) if False else None
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamSynchronize(ctypes.c_void_p(p.stream))
    assert relerr(Aq.cpu().numpy(), ref) < 1e-12


# ---------------------------------------------------------------------------
# end-to-end: the reference's known answers and iteration parity
# ---------------------------------------------------------------------------
REF = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": 10, "BOX NY": 10, "BOX NZ": 10,
       "BOX DIMX": 1, "BOX DIMY": 1, "BOX DIMZ": 1, "BOX BOUNDARY FLAG": 1, "POLYNOMIAL DEGREE": 4,
       "DATA FILE": "data/ellipticSine3D.h", "DISCRETIZATION": "CONTINUOUS", "LINEAR SOLVER": "PCG",
       "PRECONDITIONER": "NONE"}


@pytest.mark.parametrize("precon,flag,refnorm", [("NONE", 1, 0.353553390458384), ("JACOBI", 1, 0.353553400508458),
                                                 ("NONE", -1, 0.059540839002614)])
def test_run_known_answers(precon, flag, refnorm):
    import libparanumal_amd as lp
    s = dict(REF, PRECONDITIONER=precon, **{"BOX BOUNDARY FLAG": flag})
    onorm, oit, _ = oracle_lib.run(s)
    r = lp.Elliptic(s).run()
    assert abs(r["norm"] - refnorm) < 1e-5          # harness tolerance (test/test.py:63)
    assert abs(r["norm"] - onorm) < 1e-10           # against the oracle
    assert abs(r["iterations"] - oit) <= 1          # PCG iteration parity


def test_pcg_jacobi_c1_iteration_parity():
    # configuration 1: 8^3 box on [-1,1]^3, N=4, lambda=0, Dirichlet, PCG+Jacobi
    import libparanumal_amd as lp
    s = dict(REF, PRECONDITIONER="JACOBI", LAMBDA=0, **{"BOX NX": 8, "BOX NY": 8, "BOX NZ": 8, "BOX DIMX": 2,
                                                         "BOX DIMY": 2, "BOX DIMZ": 2})
    onorm, oit, ng = oracle_lib.run(s)
    assert ng == 29791
    r = lp.Elliptic(s).run()
    assert abs(r["iterations"] - oit) <= 1
    assert abs(r["norm"] - onorm) < 1e-9
    assert r["error_l2"] < 1e-3


def test_fused_jacobi_pcg_matches_unfused():
    """Jacobi-PCG in fused launches (p.Ap in the gather, z and r.z in the update)
    against the six-launch iteration: the same iterates up to the partition of the
    reductions. Also through the chunked solver (device STOP flag), which the
    p-MG coarse level uses when its N=1 kernels are off."""
    import libparanumal_amd as lp
    s = dict(REF, PRECONDITIONER="JACOBI", LAMBDA=0, **{"BOX NX": 8, "BOX NY": 8, "BOX NZ": 8, "BOX DIMX": 2,
                                                         "BOX DIMY": 2, "BOX DIMZ": 2})
    a = lp.Elliptic(dict(s, **{"PCG JACOBI FUSED": "FALSE"})).run()
    b = lp.Elliptic(dict(s, **{"PCG JACOBI FUSED": "TRUE"})).run()
    print("fused pcg", a["iterations"], b["iterations"], a["norm"], b["norm"], a["r_final"], b["r_final"], flush=True)
    assert abs(a["iterations"] - b["iterations"]) <= 1
    assert abs(a["norm"] - b["norm"]) <= 1e-10 * a["norm"]
    assert b["r_final"] <= 1e-8 * b["r0"] * (1 + 1e-12) or b["r_final"] <= 1e-8
    m = mg_settings("FP64", **{"LINEAR SOLVER STOPPING CRITERION": "RELATIVE", "MULTIGRID COARSE TOLERANCE": 1e-2,
                               "MULTIGRID COARSE MAXIT": 60, "MULTIGRID COARSE FUSED": "FALSE",
                               "MULTIGRID COARSE CHECK EVERY": 4})
    c = lp.Elliptic(dict(m, **{"PCG JACOBI FUSED": "FALSE"}))
    d = lp.Elliptic(dict(m, **{"PCG JACOBI FUSED": "TRUE"}))
    rc, rd = c.run(), d.run()
    print("fused coarse", rc["iterations"], rd["iterations"], c.geti("coarseIterations"), d.geti("coarseIterations"),
          flush=True)
    assert abs(rc["iterations"] - rd["iterations"]) <= 1
    assert abs(c.geti("coarseIterations") - d.geti("coarseIterations")) <= 1
    assert abs(rc["norm"] - rd["norm"]) <= 1e-9 * abs(rc["norm"])


# ---------------------------------------------------------------------------
# p-multigrid
# ---------------------------------------------------------------------------
def mg_settings(prec, **kw):
    s = box(3, 7, 1, 0.05, lam=1.0, PRECONDITIONER="MULTIGRID", **{"LINEAR SOLVER": "FPCG",
        "MULTIGRID COARSENING": "CUSTOM", "MULTIGRID DEGREES": "7,3,1", "MULTIGRID PRECISION": prec,
        "MULTIGRID COARSE TOLERANCE": 0.0, "MULTIGRID COARSE MAXIT": 15})
    s.update(kw)
    return s


def test_mg_vcycle_fp64_vs_oracle():
    import libparanumal_amd as lp
    s = mg_settings("FP64")
    o = OracleElliptic(s)
    o.mg_setup()
    p = lp.Elliptic(s)
    r = np.random.default_rng(11).standard_normal(o.geti("Ngather"))
    ref = o.mg_apply(r)
    rd = dev(r)
    z = torch.zeros_like(rd)
    p.precon(rd, z)
    import ctypes
    ctypes.CDLL("libamdhip64.so").hipStreamSynchronize(ctypes.c_void_p(p.stream))
    assert relerr(z.cpu().numpy(), ref) < 1e-9


def test_mg_vcycle_fp32_close_to_fp64_oracle():
    import libparanumal_amd as lp
    s = mg_settings("FP32")
    o = OracleElliptic(s)
    o.mg_setup()
    p = lp.Elliptic(s)
    r = np.random.default_rng(12).standard_normal(o.geti("Ngather"))
    ref = o.mg_apply(r)
    rd = dev(r)
    z = torch.zeros_like(rd)
    p.precon(rd, z)
    import ctypes
    ctypes.CDLL("libamdhip64.so").hipStreamSynchronize(ctypes.c_void_p(p.stream))
    assert relerr(z.cpu().numpy(), ref) < 1e-4


@pytest.mark.parametrize("prec", ["FP32", "FP64"])
def test_fpcg_pmg_solve_reaches_relative_tolerance(prec):
    import libparanumal_amd as lp
    s = mg_settings(prec, **{"LINEAR SOLVER STOPPING CRITERION": "RELATIVE", "MULTIGRID COARSE TOLERANCE": 1e-2,
                             "MULTIGRID COARSE MAXIT": 100, "BOX NX": 4})
    r = lp.Elliptic(s).run()
    assert r["r_final"] <= 1e-8 * r["r0"] * (1 + 1e-12)
    assert r["iterations"] < 40
    # same discrete solution as the all-FP64 oracle (Jacobi-PCG to a tight tolerance)
    onorm, _, _ = oracle_lib.run(dict(s, PRECONDITIONER="JACOBI", **{"LINEAR SOLVER": "PCG",
                                                                      "LINEAR SOLVER TOLERANCE": 1e-12}))
    assert abs(r["norm"] - onorm) < 1e-8 * onorm


# ---------------------------------------------------------------------------
# non-blocking solvers (linearSolverNBPCG.cpp, linearSolverNBFPCG.cpp)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("solver,precon", [("NBPCG", "JACOBI"), ("NBFPCG", "JACOBI"), ("NBPCG", "NONE")])
def test_nonblocking_solvers_vs_oracle(solver, precon):
    import libparanumal_amd as lp
    s = box(4, 4, 1, 0.05, lam=0.0, PRECONDITIONER=precon, **{"LINEAR SOLVER": solver})
    onorm, oit, _ = oracle_lib.run(s)
    r = lp.Elliptic(s).run()
    print(solver, precon, "product", r["iterations"], r["norm"], "oracle", oit, onorm, flush=True)
    assert abs(r["iterations"] - oit) <= 1
    assert abs(r["norm"] - onorm) < 1e-9 * onorm


def test_nbfpcg_pmg_fp32_reaches_tolerance():
    import libparanumal_amd as lp
    s = mg_settings("FP32", **{"LINEAR SOLVER": "NBFPCG", "LINEAR SOLVER STOPPING CRITERION": "RELATIVE",
                               "MULTIGRID COARSE TOLERANCE": 1e-2, "MULTIGRID COARSE MAXIT": 100, "BOX NX": 4})
    r = lp.Elliptic(s).run()
    f = lp.Elliptic(dict(s, **{"LINEAR SOLVER": "FPCG"})).run()
    print("NBFPCG", r["iterations"], "FPCG", f["iterations"], flush=True)
    assert r["r_final"] <= 1e-8 * r["r0"] * (1 + 1e-12)
    assert abs(r["iterations"] - f["iterations"]) <= 2
    assert abs(r["norm"] - f["norm"]) < 1e-7 * f["norm"]


# ---------------------------------------------------------------------------
# the reference harness's own setup.rc, through the ellipticMain driver
# (test/testElliptic.py:34-72 writes it; test/test.py:91-99 formats it)
# ---------------------------------------------------------------------------
def harness_rc(precon, flag=1, linear_solver="PCG", extra=()):
    keys = [("FORMAT", "2.0"), ("DATA FILE", "/path/to/solvers/elliptic/data/ellipticSine3D.h"), ("MESH FILE", "BOX"),
            ("MESH DIMENSION", 3), ("ELEMENT TYPE", 12), ("BOX NX", 10), ("BOX NY", 10), ("BOX NZ", 10),
            ("BOX DIMX", 1), ("BOX DIMY", 1), ("BOX DIMZ", 1), ("BOX BOUNDARY FLAG", flag), ("POLYNOMIAL DEGREE", 4),
            ("THREAD MODEL", "HIP"), ("PLATFORM NUMBER", 0), ("DEVICE NUMBER", 0), ("DISCRETIZATION", "CONTINUOUS"),
            ("LINEAR SOLVER", linear_solver), ("PRECONDITIONER", precon), ("MULTIGRID SMOOTHER", "CHEBYSHEV"),
            ("PARALMOND CYCLE", "VCYCLE"), ("PARALMOND STRENGTH", "SYMMETRIC"), ("PARALMOND AGGREGATION", "UNSMOOTHED"),
            ("PARALMOND SMOOTHER", "CHEBYSHEV"), ("OUTPUT TO FILE", "FALSE"), ("VERBOSE", "FALSE")] + list(extra)
    return "".join(f"[{k}]\n{v}\n\n" for k, v in keys)


@pytest.mark.parametrize("precon,flag,solver,refnorm", [
    ("NONE", 1, "PCG", 0.353553390458384),       # testEllipticHex_C0 (testElliptic.py:95-98)
    ("JACOBI", 1, "PCG", 0.353553400508458),     # testEllipticHex_C0_Jacobi (:199-203)
    ("MULTIGRID", 1, "FPCG", 0.353553400508458), # testEllipticHex_C0_Multigrid (:209-213), coarse level differs
    ("NONE", -1, "PCG", 0.059540839002614),      # testEllipticHex_C0_AllNeumann (:244-248)
    ("MULTIGRID-AMG", 1, "FPCG", 0.353553400508458),  # testEllipticHex_C0_Multigrid with the parAlmond base level
])
def test_elliptic_main_with_harness_settings(tmp_path, precon, flag, solver, refnorm):
    import subprocess
    import libparanumal_amd as lp
    rc = tmp_path / "setup.rc"
    extra = ()
    if precon == "MULTIGRID-AMG":
        precon, extra = "MULTIGRID", (("MULTIGRID COARSE SOLVER", "AMG"),)
    rc.write_text(harness_rc(precon, flag, solver, extra))
    run = subprocess.run([lp.MAIN_PATH, str(rc)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert run.returncode == 0, run.stderr[-2000:]
    last = run.stdout.strip().splitlines()[-1]
    assert last.startswith("Solution norm = "), last
    norm = float(last.split("=")[1])
    print(precon, flag, solver, norm, flush=True)
    assert abs(norm - refnorm) < 1e-5  # the harness tolerance (test/test.py)


def test_coarse_graph_replays_match_eager():
    """The fused coarse Jacobi-PCG replays captured HIP graphs; every replay (several
    solves with eager work and host reads in between) must give the eager result."""
    import libparanumal_amd as lp
    base = mg_settings("FP32", **{"LINEAR SOLVER STOPPING CRITERION": "RELATIVE", "MULTIGRID COARSE TOLERANCE": 1e-2,
                                  "MULTIGRID COARSE MAXIT": 100, "BOX NX": 4})
    eager = lp.Elliptic(dict(base, **{"MULTIGRID COARSE GRAPH": "FALSE"}))
    graph = lp.Elliptic(dict(base, **{"MULTIGRID COARSE GRAPH": "TRUE"}))
    for rep in range(3):
        a, b = eager.run(), graph.run()
        print("replay", rep, a["iterations"], b["iterations"], a["norm"], b["norm"], flush=True)
        assert a["iterations"] == b["iterations"]
        assert a["norm"] == b["norm"]
        assert eager.geti("coarseIterations") == graph.geti("coarseIterations")
        assert np.array_equal(eager.solution(), graph.solution())


def test_coarse_single_reduction_matches_three_kernel():
    """The coarse Jacobi-PCG in its single-reduction form (Chronopoulos-Gear,
    two launches per iteration) against the three-launch form: with a fixed
    count of coarse iterations (tolerance 0) the p-MG solves agree to rounding
    and take the same outer iterations; with a tolerance the coarse counts
    agree to +-1. Repeated solves (graph replays, parity-alternating buffers)
    are bit-identical."""
    import libparanumal_amd as lp
    for ctol in (0.0, 1e-2):
        base = mg_settings("FP32", **{"LINEAR SOLVER STOPPING CRITERION": "RELATIVE", "BOX NX": 4,
                                      "MULTIGRID COARSE TOLERANCE": ctol, "MULTIGRID COARSE MAXIT": 40})
        three = lp.Elliptic(dict(base, **{"MULTIGRID COARSE SINGLE REDUCTION": "FALSE"}))
        one = lp.Elliptic(dict(base, **{"MULTIGRID COARSE SINGLE REDUCTION": "TRUE"}))
        a = three.run()
        runs = [one.run() for _ in range(3)]
        b = runs[0]
        print("single reduction", ctol, a["iterations"], b["iterations"], a["norm"], b["norm"],
              three.geti("coarseIterations"), one.geti("coarseIterations"), flush=True)
        assert all(r["norm"] == b["norm"] and r["iterations"] == b["iterations"] for r in runs)
        assert abs(a["iterations"] - b["iterations"]) <= (0 if ctol == 0 else 1)
        assert abs(a["norm"] - b["norm"]) <= 1e-9 * abs(a["norm"])
        assert b["r_final"] <= 1e-8 * b["r0"] * (1 + 1e-12)
        if ctol == 0:
            assert one.geti("coarseIterations") == three.geti("coarseIterations") == 40
            x1, x3 = one.solution(), three.solution()
            assert np.max(np.abs(x1 - x3)) <= 1e-9 * np.max(np.abs(x3))
        else:
            assert abs(one.geti("coarseIterations") - three.geti("coarseIterations")) <= 1

def test_coarse_assembled_matches_element_form():
    """The single-reduction coarse Jacobi-PCG on the assembled N=1 operator
    (one launch per iteration) against its element form: with a fixed count
    of coarse iterations the p-MG solves agree to rounding and take the same
    outer iterations; with a tolerance the coarse counts agree to +-1.
    Repeated solves are bit-identical. The 52^3-element case has 132,651
    coarse rows, more than one pass of the kernel's 1024 blocks of 128 rows."""
    import libparanumal_amd as lp
    cases = [(0.0, {"BOX NX": 4}), (1e-2, {"BOX NX": 4}),
             (1e-2, {"BOX NX": 52, "BOX NY": 52, "BOX NZ": 52, "POLYNOMIAL DEGREE": 3,
                                       "MULTIGRID DEGREES": "3,1"})]
    for ctol, kw in cases:
        base = mg_settings("FP32", **{"LINEAR SOLVER STOPPING CRITERION": "RELATIVE",
                                      "MULTIGRID COARSE TOLERANCE": ctol, "MULTIGRID COARSE MAXIT": 40}, **kw)
        elem = lp.Elliptic(dict(base, **{"MULTIGRID COARSE ASSEMBLED": "FALSE"}))
        asm = lp.Elliptic(dict(base, **{"MULTIGRID COARSE ASSEMBLED": "TRUE"}))
        a = elem.run()
        runs = [asm.run() for _ in range(3)]
        b = runs[0]
        print("assembled coarse", ctol, kw.get("BOX NX"), a["iterations"], b["iterations"], a["norm"], b["norm"],
              elem.geti("coarseIterations"), asm.geti("coarseIterations"), flush=True)
        assert all(r["norm"] == b["norm"] and r["iterations"] == b["iterations"] for r in runs)
        assert abs(a["iterations"] - b["iterations"]) <= (0 if ctol == 0 else 1)
        assert abs(a["norm"] - b["norm"]) <= 1e-9 * abs(a["norm"])
        assert b["r_final"] <= 1e-8 * b["r0"] * (1 + 1e-12)
        if ctol == 0:
            assert asm.geti("coarseIterations") == elem.geti("coarseIterations") == 40
            x1, x3 = asm.solution(), elem.solution()
            assert np.max(np.abs(x1 - x3)) <= 1e-9 * np.max(np.abs(x3))
        else:
            assert abs(asm.geti("coarseIterations") - elem.geti("coarseIterations")) <= 1


def test_coarse_assembled_vcycle_vs_oracle():
    """The FP64 V-cycle with the assembled single-reduction coarse solve (15
    coarse iterations) against the oracle's FP64 V-cycle."""
    import ctypes
    import libparanumal_amd as lp
    s = mg_settings("FP64", **{"MULTIGRID COARSE ASSEMBLED": "TRUE"})
    o = OracleElliptic(s)
    o.mg_setup()
    p = lp.Elliptic(s)
    r = np.random.default_rng(17).standard_normal(o.geti("Ngather"))
    ref = o.mg_apply(r)
    rd = dev(r)
    z = torch.zeros_like(rd)
    p.precon(rd, z)
    ctypes.CDLL("libamdhip64.so").hipStreamSynchronize(ctypes.c_void_p(p.stream))
    assert relerr(z.cpu().numpy(), ref) < 1e-9


def test_vtu_output_matches_manufactured_solution(tmp_path):
    """[OUTPUT TO FILE] TRUE writes <OUTPUT FILE NAME>_0000.vtu in the reference's
    layout (ellipticPlotFields.cpp): (N+4)^3 equispaced plot nodes per element,
    6 (N+3)^3 tetrahedra; the plotted field is the solution, which for the
    sine data is close to sin(pi x) sin(pi y) sin(pi z) at every plot node."""
    import xml.etree.ElementTree as ET

    import libparanumal_amd as lp
    s = dict(REF, **{"BOX NX": 4, "BOX NY": 4, "BOX NZ": 4, "POLYNOMIAL DEGREE": 5, "PRECONDITIONER": "JACOBI",
                     "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "OUTPUT TO FILE": "TRUE",
                     "OUTPUT FILE NAME": str(tmp_path / "sol"), "LINEAR SOLVER STOPPING CRITERION": "RELATIVE",
                     "LINEAR SOLVER TOLERANCE": 1e-10})
    el = lp.Elliptic(s)
    r = el.run()
    f = tmp_path / "sol_0000.vtu"
    assert f.exists()
    piece = ET.parse(f).getroot().find("UnstructuredGrid").find("Piece")
    E, pNq = 64, 5 + 4
    assert int(piece.get("NumberOfPoints")) == E * pNq ** 3
    assert int(piece.get("NumberOfCells")) == E * 6 * (pNq - 1) ** 3
    pts = np.array(piece.find("Points").find("DataArray").text.split(), dtype=float).reshape(-1, 3)
    fld = np.array(piece.find("PointData").find("DataArray").text.split(), dtype=float)
    conn = np.array(piece.find("Cells").findall("DataArray")[0].text.split(), dtype=np.int64)
    assert conn.size == 4 * E * 6 * (pNq - 1) ** 3 and conn.min() == 0 and conn.max() == E * pNq ** 3 - 1
    assert np.abs(pts).max() <= 1.0 + 1e-6
    exact = np.sin(np.pi * pts[:, 0]) * np.sin(np.pi * pts[:, 1]) * np.sin(np.pi * pts[:, 2])
    err = np.abs(fld - exact).max()
    print("vtu points", len(pts), "max |u_h - u| at plot nodes", err, r["iterations"], flush=True)
    assert err < 1e-3


# ---------------------------------------------------------------------------
# Gmsh mesh files
# ---------------------------------------------------------------------------
def msh_settings(N, **kw):
    import os
    root = os.path.dirname(os.path.abspath(__file__))
    s = {"MESH FILE": os.path.join(root, "golden", "cubeHex.msh"), "MESH DIMENSION": 3, "ELEMENT TYPE": 12,
         "POLYNOMIAL DEGREE": N, "LAMBDA": 1.0, "DATA FILE": "data/ellipticSine3D.h",
         "DISCRETIZATION": "CONTINUOUS", "LINEAR SOLVER": "PCG", "PRECONDITIONER": "JACOBI",
         "LINEAR SOLVER STOPPING CRITERION": "RELATIVE", "LINEAR SOLVER TOLERANCE": 1e-10}
    s.update(kw)
    return s


@pytest.mark.parametrize("N,precon", [(3, "JACOBI"), (4, "MULTIGRID")])
def test_gmsh_cube_matches_box(N, precon):
    """The reference's test mesh test/cubeHex.msh (read by the Gmsh reader, nodes
    numbered by the general path) is the uniform 6^3 grid of [-1,1]^3 with every
    boundary face tagged 1: it must give the BOX solve's dof count and solution
    norm (element and node orders differ, so the sums differ in rounding only)."""
    import libparanumal_amd as lp
    extra = {}
    if precon == "MULTIGRID":
        extra = {"LINEAR SOLVER": "FPCG", "MULTIGRID COARSENING": "CUSTOM", "MULTIGRID DEGREES": f"{N},1",
                 "MULTIGRID PRECISION": "FP32"}
    m = lp.Elliptic(msh_settings(N, PRECONDITIONER=precon, **extra))
    b = lp.Elliptic(dict(msh_settings(N, PRECONDITIONER=precon, **extra), **{
        "MESH FILE": "BOX", "BOX NX": 6, "BOX NY": 6, "BOX NZ": 6, "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2,
        "BOX BOUNDARY FLAG": 1}))
    assert m.geti("Nelements") == 216 and m.geti("Ngather") == b.geti("Ngather") == (6 * N - 1) ** 3
    rm, rb = m.run(), b.run()
    print("gmsh", N, precon, rm["iterations"], rb["iterations"], rm["norm"], rb["norm"], flush=True)
    assert rm["r_final"] <= 1e-10 * rm["r0"] * (1 + 1e-12)
    assert abs(rm["norm"] - rb["norm"]) <= 1e-9 * rb["norm"]
    assert abs(rm["iterations"] - rb["iterations"]) <= 1


def test_gmsh_neumann_tags_match_box(tmp_path):
    """cubeHex.msh with its boundary faces retagged 2 (Neumann, written to a
    temporary copy) against the BOX mesh with BOX BOUNDARY FLAG 2, screened
    (lambda = 1) so that the problem is definite: every node is a dof and the
    solutions agree."""
    import os
    import libparanumal_amd as lp
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cubeHex.msh")
    lines = open(src).read().split("\n")
    out, inside = [], False
    for ln in lines:
        if ln.startswith("$Elements"):
            inside = True
        elif ln.startswith("$EndElements"):
            inside = False
        t = ln.split()
        if inside and len(t) > 3 and t[1] == "3":  # quadrilateral: id type ntags tag...
            t[3] = "2"
            ln = " ".join(t)
        out.append(ln)
    msh = tmp_path / "cubeHexNeumann.msh"
    msh.write_text("\n".join(out))
    m = lp.Elliptic(msh_settings(3, **{"MESH FILE": str(msh), "LAMBDA": 1.0}))
    b = lp.Elliptic(dict(msh_settings(3, LAMBDA=1.0), **{
        "MESH FILE": "BOX", "BOX NX": 6, "BOX NY": 6, "BOX NZ": 6, "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2,
        "BOX BOUNDARY FLAG": 2}))
    assert m.geti("Ngather") == b.geti("Ngather") == 19 ** 3
    rm, rb = m.run(), b.run()
    print("gmsh neumann", rm["iterations"], rb["iterations"], rm["norm"], rb["norm"], flush=True)
    assert abs(rm["norm"] - rb["norm"]) <= 1e-9 * rb["norm"]

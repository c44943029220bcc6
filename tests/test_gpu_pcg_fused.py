"""The fused Jacobi-PCG iteration against the oracle and the separate kernels.

After the p update, libp_pcg_ax_hex_pap_f64 runs the partial Ax of p and forms
p.Ap from the local nodes and alpha; libp_pcg_gather_jacobi_update_f64 gathers
A p row by row and applies the update of linearSolverPCG.cpp:142-160 with the
Jacobi preconditioner (ellipticPreconJacobi.cpp:42-51) without storing Ap.
Checked through the C ABI against the oracle's partial Ax and operator
(solvers/elliptic/okl/ellipticAxHex3D.okl:156-295, ellipticOperator.cpp:29-107)
and, bit for bit, against the separate kernels (libp_ax_hex_partial_f64, or the
FP64 MFMA kernel libp_ax_hex_partial_mfma_n15 at N=15, libp_ogs_gather_f64 +
libp_pcg_jacobi_update_f64); then whole solves with and without the fused form."""
import numpy as np
import pytest

from tests.oracle_lib import OracleElliptic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def lp():
    import libparanumal_amd
    return libparanumal_amd


def box(nx, N, curve=0.05, lam=0.5, **kw):
    s = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": nx, "BOX NY": nx + 1, "BOX NZ": nx,
         "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "BOX BOUNDARY FLAG": 1, "POLYNOMIAL DEGREE": N,
         "BOX CURVED MAP AMPLITUDE": curve, "LAMBDA": lam, "DATA FILE": "data/ellipticSine3D.h",
         "DISCRETIZATION": "CONTINUOUS", "LINEAR SOLVER": "PCG", "PRECONDITIONER": "JACOBI"}
    s.update(kw)
    return s


@pytest.mark.parametrize("N", [1, 2, 3, 7, 15])
def test_fused_iteration_through_c_abi(N):
    o = OracleElliptic(box(3 if N < 15 else 2, N))
    L = lp().lib()
    E, Np, Ng = o.geti("Nelements"), o.geti("Np"), o.geti("Ngather")
    NL = E * Np
    rs, ci = o.array("rowStartsT"), o.array("colIdsT")
    assert rs.size == Ng + 1
    rng = np.random.default_rng(10 * N)
    p = rng.uniform(-1, 1, Ng)
    x0 = rng.uniform(-1, 1, Ng)
    r0 = rng.uniform(-1, 1, Ng)
    invD = 1.0 / o.diagonal()
    rdotz = 1.9
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    g2l, wJ, ggeo, D = dev(o.array("GlobalToLocal")), dev(o.array("wJ")), dev(o.array("ggeo")), dev(o.array("D"))
    pd, invd, rsd, cid = dev(p), dev(invD), dev(rs), dev(ci)
    AqL = torch.full((NL,), float("nan"), dtype=torch.float64, device="cuda")
    part = torch.zeros(L.libp_ax_hex_split_blocks(N, E), dtype=torch.float64, device="cuda")
    s = torch.zeros(16, dtype=torch.float64, device="cuda")
    s[0] = rdotz
    st = torch.cuda.current_stream().cuda_stream
    lp().check(L.libp_pcg_ax_hex_pap_f64(N, E, None, g2l.data_ptr(), wJ.data_ptr(), ggeo.data_ptr(), D.data_ptr(),
                                         0.5, pd.data_ptr(), AqL.data_ptr(), part.data_ptr(), s.data_ptr(), st))
    AqL2 = torch.full((NL,), float("nan"), dtype=torch.float64, device="cuda")
    if N == 15:  # the FP64 MFMA kernel, as the operator runs it at N=15
        lp().check(L.libp_ax_hex_partial_mfma_n15(E, None, g2l.data_ptr(), wJ.data_ptr(), ggeo.data_ptr(),
                                                  D.data_ptr(), 0.5, pd.data_ptr(), AqL2.data_ptr(), st))
    else:
        lp().check(L.libp_ax_hex_partial_f64(N, E, None, g2l.data_ptr(), wJ.data_ptr(), ggeo.data_ptr(),
                                             D.data_ptr(), 0.5, pd.data_ptr(), AqL2.data_ptr(), st))
    torch.cuda.synchronize()
    # the same E-vector as the plain partial Ax, bit for bit, and the oracle's to 1e-12
    assert torch.equal(AqL, AqL2)
    AqL_ref = o.partial_ax(p)
    assert np.max(np.abs(AqL.cpu().numpy() - AqL_ref)) <= 1e-12 * np.max(np.abs(AqL_ref))
    Ap = o.operator(p)
    pap = float(p @ Ap)
    sh = s.cpu().numpy()
    print(f"N={N}: p.Ap {sh[4]:.17e} oracle {pap:.17e}", flush=True)
    assert abs(sh[4] - pap) <= 1e-12 * abs(pap)
    assert abs(sh[2] - rdotz / sh[4]) <= 1e-15 * abs(sh[2])

    # the fused gather + update against numpy, and bit for bit against the separate kernels
    ws = torch.empty(L.libp_reduce_ws_bytes(), dtype=torch.uint8, device="cuda")
    gpart = torch.zeros(L.libp_pcg_gather_jacobi_partials(Ng), dtype=torch.float64, device="cuda")
    xa, ra, za = dev(x0), dev(r0), torch.full((Ng,), float("nan"), dtype=torch.float64, device="cuda")
    sa = s.clone()
    lp().check(L.libp_pcg_gather_jacobi_update_f64(Ng, rsd.data_ptr(), cid.data_ptr(), AqL.data_ptr(),
                                                   invd.data_ptr(), pd.data_ptr(), xa.data_ptr(), ra.data_ptr(),
                                                   za.data_ptr(), sa.data_ptr(), gpart.data_ptr(), st))
    Apd = torch.empty(Ng, dtype=torch.float64, device="cuda")
    lp().check(L.libp_ogs_gather_f64(Ng, rsd.data_ptr(), cid.data_ptr(), AqL.data_ptr(), Apd.data_ptr(), st))
    xb, rb, zb = dev(x0), dev(r0), torch.full((Ng,), float("nan"), dtype=torch.float64, device="cuda")
    sb = s.clone()
    lp().check(L.libp_pcg_jacobi_update_f64(Ng, invd.data_ptr(), pd.data_ptr(), Apd.data_ptr(), xb.data_ptr(),
                                            rb.data_ptr(), zb.data_ptr(), sb.data_ptr(), ws.data_ptr(), st))
    torch.cuda.synchronize()
    assert torch.equal(xa, xb) and torch.equal(ra, rb) and torch.equal(za, zb)
    alpha = sh[2]
    x_ref, r_ref = x0 + alpha * p, r0 - alpha * Ap
    assert np.max(np.abs(xa.cpu().numpy() - x_ref)) <= 4e-15 * np.max(np.abs(x_ref))
    assert np.max(np.abs(ra.cpu().numpy() - r_ref)) <= 1e-12 * np.max(np.abs(r_ref))
    assert np.max(np.abs(za.cpu().numpy() - invD * r_ref)) <= 1e-12 * np.max(np.abs(invD * r_ref))
    sah, sbh = sa.cpu().numpy(), sb.cpu().numpy()
    rr, rz = float(r_ref @ r_ref), float(r_ref @ (invD * r_ref))
    assert abs(sah[5] - rr) <= 1e-11 * rr and abs(sah[0] - rz) <= 1e-11 * abs(rz)
    assert abs(sah[5] - sbh[5]) <= 1e-13 * rr and abs(sah[0] - sbh[0]) <= 1e-13 * abs(rz)
    assert sah[1] == rdotz and sah[9] == 1.0 and sah[8] == 0.0
    assert abs(sah[3] - sah[0] / rdotz) <= 1e-15 * abs(sah[3])


@pytest.mark.parametrize("N", [3, 15])
def test_stop_flag_skips_both_kernels(N):
    """With scalars[STOP] set, neither kernel writes anything (chunked solves queue
    iterations past convergence)."""
    o = OracleElliptic(box(3 if N < 15 else 2, N))
    L = lp().lib()
    E, Np, Ng = o.geti("Nelements"), o.geti("Np"), o.geti("Ngather")
    rs, ci = o.array("rowStartsT"), o.array("colIdsT")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    g2l, wJ, ggeo, D = dev(o.array("GlobalToLocal")), dev(o.array("wJ")), dev(o.array("ggeo")), dev(o.array("D"))
    z = torch.ones(Ng, dtype=torch.float64, device="cuda")
    AqL = torch.full((E * Np,), 7.0, dtype=torch.float64, device="cuda")
    part = torch.zeros(L.libp_ax_hex_split_blocks(N, E), dtype=torch.float64, device="cuda")
    s = torch.zeros(16, dtype=torch.float64, device="cuda")
    s[8] = 1.0
    s0 = s.clone()
    st = torch.cuda.current_stream().cuda_stream
    lp().check(L.libp_pcg_ax_hex_pap_f64(N, E, None, g2l.data_ptr(), wJ.data_ptr(), ggeo.data_ptr(), D.data_ptr(),
                                         0.0, z.data_ptr(), AqL.data_ptr(), part.data_ptr(), s.data_ptr(), st))
    gpart = torch.zeros(L.libp_pcg_gather_jacobi_partials(Ng), dtype=torch.float64, device="cuda")
    x, r, zz = z.clone(), z.clone(), z.clone()
    lp().check(L.libp_pcg_gather_jacobi_update_f64(Ng, dev(rs).data_ptr(), dev(ci).data_ptr(), AqL.data_ptr(),
                                                   z.data_ptr(), z.data_ptr(), x.data_ptr(), r.data_ptr(),
                                                   zz.data_ptr(), s.data_ptr(), gpart.data_ptr(), st))
    torch.cuda.synchronize()
    assert bool((AqL == 7.0).all())
    assert bool((x == 1.0).all()) and bool((r == 1.0).all()) and bool((zz == 1.0).all())
    assert torch.equal(s, s0)


def test_stop_flag_coarse_iterations_write_nothing():
    """Coarse Jacobi-PCG launches queued after STOP (the chunks run ahead of the host's
    check) write nothing: the single-reduction iteration on a 70,000-row CSR matrix
    (547 blocks) in every mode and parity, and the three-kernel N=1 iteration. Each
    block reads the flag once and all its waves return together."""
    import ctypes
    L = lp().lib()
    st = torch.cuda.current_stream().cuda_stream
    f64 = dict(dtype=torch.float64, device="cuda")
    n = 70000
    rows = np.arange(n)
    starts = np.concatenate([[0], np.cumsum(np.where((rows == 0) | (rows == n - 1), 2, 3))]).astype(np.int32)
    cols = np.concatenate([[j for j in (i - 1, i, i + 1) if 0 <= j < n] for i in range(n)]).astype(np.int32)
    vals = np.concatenate([[-1.0 if j != i else 2.0 for j in (i - 1, i, i + 1) if 0 <= j < n] for i in range(n)])
    rs, ci, va = (torch.from_numpy(a).cuda() for a in (starts, cols, vals))
    invD = torch.full((n,), 0.5, **f64)
    vec = lambda: torch.full((n,), 7.0, **f64)
    x = vec()
    pairs = {k: [vec(), vec()] for k in "wpsru"}
    arr = {k: (ctypes.c_void_p * 2)(*(t.data_ptr() for t in v)) for k, v in pairs.items()}
    wsb = L.libp_reduce_ws_bytes()
    ws = torch.full((2 * wsb // 8 + 1,), 3.0, **f64)
    s = torch.zeros(16, **f64)
    s[7], s[8] = 1e-3, 1.0
    s0, ws0 = s.clone(), ws.clone()
    for mode in (0, 1, 2):
        for parity in (0, 1):
            lp().check(L.libp_pcg_jacobi_csr_cg1_f64(n, rs.data_ptr(), ci.data_ptr(), va.data_ptr(), invD.data_ptr(),
                                                     1e-2, mode, parity, x.data_ptr(), arr["w"], arr["p"], arr["s"],
                                                     arr["r"], arr["u"], s.data_ptr(), ws.data_ptr(), st))
    torch.cuda.synchronize()
    assert bool((x == 7.0).all())
    assert all(bool((t == 7.0).all()) for v in pairs.values() for t in v)
    assert torch.equal(s, s0) and torch.equal(ws, ws0)

    o = OracleElliptic(box(5, 1))
    E, Ng = o.geti("Nelements"), o.geti("Ngather")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    g2l, wJ, ggeo, D = dev(o.array("GlobalToLocal")), dev(o.array("wJ")), dev(o.array("ggeo")), dev(o.array("D"))
    grs, gci = dev(o.array("rowStartsT")), dev(o.array("colIdsT"))
    v1 = [torch.full((Ng,), 7.0, **f64) for _ in range(6)]  # invD, x, r, z, p, Ap
    AqL = torch.full((E * 8,), 7.0, **f64)
    for first in (1, 0):
        lp().check(L.libp_pcg_jacobi_hex1_iteration_f64(E, g2l.data_ptr(), wJ.data_ptr(), ggeo.data_ptr(),
                                                        D.data_ptr(), 1.0, Ng, grs.data_ptr(), gci.data_ptr(),
                                                        *[t.data_ptr() for t in v1[:1]], first,
                                                        *[t.data_ptr() for t in v1[1:]], AqL.data_ptr(),
                                                        s.data_ptr(), ws.data_ptr(), st))
    torch.cuda.synchronize()
    assert all(bool((t == 7.0).all()) for t in v1) and bool((AqL == 7.0).all())
    assert torch.equal(s, s0) and torch.equal(ws, ws0)


def test_fused_kernels_reject_bad_arguments():
    L = lp().lib()
    v = torch.zeros(64, dtype=torch.float64, device="cuda")
    ptr = v.data_ptr()
    assert L.libp_pcg_ax_hex_pap_f64(16, 1, None, ptr, ptr, ptr, ptr, 0.0, ptr, ptr, ptr, ptr, None) == 1
    assert b"degree" in L.libp_last_error()
    assert L.libp_ax_hex_split_blocks(15, 7) == 7 and L.libp_ax_hex_split_blocks(16, 7) == -1
    assert L.libp_pcg_gather_jacobi_update_f64(-1, ptr, ptr, ptr, ptr, ptr, ptr, ptr, ptr, ptr, ptr, None) == 1
    assert L.libp_pcg_gather_jacobi_update_f64(8, ptr, ptr, ptr, ptr, ptr, ptr, ptr, ptr, ptr, None, None) == 1
    assert L.libp_pcg_gather_jacobi_partials(-1) == -1


@pytest.mark.parametrize("N,nx,curve,lam", [(4, 5, 0.0, 0.0), (7, 4, 0.05, 1.0), (3, 6, 0.05, 0.0),
                                            (15, 2, 0.05, 1.0)])
@pytest.mark.parametrize("stop", ["ABS/REL-INITRESID", "RELATIVE"])
def test_fused_solve_matches_four_launch_form_and_oracle(N, nx, curve, lam, stop):
    """Whole PCG+Jacobi solves (elliptic_t::Run): the fused iteration against the
    four-launch form ([PCG FUSED OPERATOR] FALSE) and the oracle: iterations within
    one, solution norms to 1e-10."""
    from tests import oracle_lib
    s = box(nx, N, curve=curve, lam=lam, **{"LINEAR SOLVER STOPPING CRITERION": stop})
    a = lp().Elliptic(dict(s, **{"PCG FUSED OPERATOR": "TRUE"}))
    ra = a.run()
    a.close()
    b = lp().Elliptic(dict(s, **{"PCG FUSED OPERATOR": "FALSE"}))
    rb = b.run()
    b.close()
    onorm, oit, _ = oracle_lib.run(s)
    print(f"N={N} {nx}^3 {stop}: fused {ra['iterations']} its {ra['norm']:.15f}; four-launch {rb['iterations']} "
          f"its {rb['norm']:.15f}; oracle {oit} its {onorm:.15f}", flush=True)
    assert abs(ra["iterations"] - rb["iterations"]) <= 1 and abs(ra["iterations"] - oit) <= 1
    assert abs(ra["norm"] - rb["norm"]) <= 1e-10 * abs(rb["norm"])
    assert abs(ra["norm"] - onorm) <= 1e-10 * abs(onorm)


def test_scalars_to_host_through_mapped_memory():
    """libp_scalars_to_host_f64 writes device scalars into pinned host memory mapped
    for the device (the host read of the solver scalars, linearSolverPCG.cpp:163):
    after an event, the host sees the values bit for bit; n outside [0, 64] is
    rejected."""
    import ctypes
    L = lp().lib()
    hip = ctypes.CDLL("libamdhip64.so")
    host, dptr = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipHostMalloc(ctypes.byref(host), ctypes.c_size_t(64 * 8), ctypes.c_uint(0x2 | 0x40000000)) == 0
    try:
        assert hip.hipHostGetDevicePointer(ctypes.byref(dptr), host, ctypes.c_uint(0)) == 0
        buf = (ctypes.c_double * 64).from_address(host.value)
        for i in range(64):
            buf[i] = -1.0
        src = torch.from_numpy(np.random.default_rng(3).standard_normal(64)).cuda()
        st = torch.cuda.current_stream()
        for n in (1, 16, 64):
            lp().check(L.libp_scalars_to_host_f64(n, src.data_ptr(), dptr.value, st.cuda_stream))
            ev = torch.cuda.Event()
            ev.record(st)
            ev.synchronize()
            got = np.array(buf[:n])
            assert np.array_equal(got, src[:n].cpu().numpy())
        assert L.libp_scalars_to_host_f64(65, src.data_ptr(), dptr.value, st.cuda_stream) == 1
        assert L.libp_scalars_to_host_f64(-1, src.data_ptr(), dptr.value, st.cuda_stream) == 1
        assert L.libp_scalars_to_host_f64(0, None, None, st.cuda_stream) == 0
    finally:
        hip.hipHostFree(host)

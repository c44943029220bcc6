"""The split form of the C0 operator (libp_ax_hex_split_f64 + libp_ogs_gather_face_f64)
against the two-pass form (partial Ax, then the full gather) and the oracle.

The operator is elliptic_t::Operator (solvers/elliptic/src/ellipticOperator.cpp:29-107).
Interior rows come straight from the Ax kernel, boundary rows from the face gather in
the reference's summation order, so the two forms must agree bit for bit."""
import ctypes

import numpy as np
import pytest

from tests.oracle_lib import OracleElliptic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def box(nx, N, flag=1, curve=0.05, lam=0.7, **kw):
    s = {"MESH FILE": "BOX", "MESH DIMENSION": 3, "ELEMENT TYPE": 12, "BOX NX": nx, "BOX NY": nx + 1,
         "BOX NZ": max(nx - 1, 1), "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2, "BOX BOUNDARY FLAG": flag,
         "POLYNOMIAL DEGREE": N, "LAMBDA": lam, "DATA FILE": "data/ellipticSine3D.h",
         "BOX CURVED MAP AMPLITUDE": curve}
    s.update(kw)
    return s


def sync(el):
    torch.cuda.synchronize()
    ctypes.CDLL("libamdhip64.so").hipStreamSynchronize(ctypes.c_void_p(el.stream))


def lp():
    import libparanumal_amd
    return libparanumal_amd


@pytest.mark.parametrize("N", list(range(1, 15)))
@pytest.mark.parametrize("flag,layout", [(1, "SLOT"), (2, "SLOT"), (1, "ELEMENT"), (2, "ELEMENT")])
def test_split_operator_bit_identical_to_two_pass(N, flag, layout):
    nx = 3 if N <= 8 else 2
    s = box(nx, N, flag=flag)
    a = lp().Elliptic(dict(s, **{"OPERATOR SPLIT GATHER": "TRUE", "OPERATOR FACE LAYOUT": layout}))
    assert a.geti("splitOperator") == 1
    b = lp().Elliptic(dict(s, **{"OPERATOR SPLIT GATHER": "FALSE"}))
    n = a.geti("Ngather")
    q = torch.rand(n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(N)) - 0.5
    ya = torch.full_like(q, float("nan"))
    yb = torch.full_like(q, float("nan"))
    a.operator(q, ya)
    b.operator(q, yb)
    sync(a)
    sync(b)
    assert torch.equal(ya, yb), f"N={N}: max diff {float((ya - yb).abs().max())}"


def face_maps(o):
    """Face index of each node (-1 interior) and the split maps, built from the
    oracle's CSR independently of the library."""
    N, Np = o.geti("N"), o.geti("Np")
    Nq = N + 1
    i, j, k = np.meshgrid(np.arange(Nq), np.arange(Nq), np.arange(Nq), indexing="ij")
    inner = ((i > 0) & (i < N) & (j > 0) & (j < N) & (k > 0) & (k < N))
    n = (i + j * Nq + k * Nq * Nq).ravel()
    interior = np.zeros(Np, bool)
    interior[n] = inner.ravel()
    fidx = -np.ones(Np, np.int64)
    fidx[~interior] = np.arange((~interior).sum())
    rs, ci = o.array("rowStartsT"), o.array("colIdsT")
    rows, starts, cols = [], [0], []
    Nface = int((~interior).sum())
    for r in range(rs.size - 1):
        c = ci[rs[r]:rs[r + 1]]
        if c.size == 1 and interior[c[0] % Np]:
            continue
        assert np.all(~interior[c % Np])
        rows.append(r)
        cols.extend((c // Np) * Nface + fidx[c % Np])
        starts.append(len(cols))
    return interior, fidx, Nface, np.array(rows, np.int32), np.array(starts, np.int32), np.array(cols, np.int32)


@pytest.mark.parametrize("N", [1, 3, 7])
@pytest.mark.parametrize("layout", ["ELEMENT", "SLOT"])
def test_split_kernels_through_c_abi_vs_oracle(N, layout):
    """libp_ax_hex_split_f64 / libp_ogs_gather_face_f64 / libp_pcg_gather_face_pap_f64
    called directly with the oracle's arrays and maps built here."""
    o = OracleElliptic(box(3, N))
    L = lp().lib()
    E, Np, Ng = o.geti("Nelements"), o.geti("Np"), o.geti("Ngather")
    interior, fidx, Nface, rows, starts, cols = face_maps(o)
    assert L.libp_ax_hex_face_nodes(N) == Nface
    rng = np.random.default_rng(N)
    q = rng.uniform(-1, 1, Ng)
    AqL = o.partial_ax(q)
    ref = o.operator(q)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    g2l, wJ, ggeo, D = dev(o.array("GlobalToLocal")), dev(o.array("wJ")), dev(o.array("ggeo")), dev(o.array("D"))
    qd = dev(q)
    Aq = torch.full((Ng,), float("nan"), dtype=torch.float64, device="cuda")
    ncop = cols.size
    if layout == "SLOT":  # each boundary node's place in the row-ordered copy list; masked nodes -> trash slot
        slots = np.full(E * Nface, ncop, np.int32)
        slots[cols] = np.arange(ncop, dtype=np.int32)
        sl = dev(slots)
        AqF = torch.full((ncop + 1,), float("nan"), dtype=torch.float64, device="cuda")
    else:
        sl = None
        AqF = torch.full((E * Nface,), float("nan"), dtype=torch.float64, device="cuda")
    nax = L.libp_ax_hex_split_blocks(N, E)
    nf = L.libp_ogs_gather_face_blocks(rows.size)
    part = torch.zeros(nax + nf, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    lp().check(L.libp_ax_hex_split_f64(N, E, None, g2l.data_ptr(), wJ.data_ptr(), ggeo.data_ptr(), D.data_ptr(), 0.7,
                                       qd.data_ptr(), Aq.data_ptr(), AqF.data_ptr(),
                                       sl.data_ptr() if sl is not None else None, part.data_ptr(), st))
    torch.cuda.synchronize()
    # face E-vector = the oracle's E-vector at the boundary nodes (in row order for the slot layout)
    AqLe = AqL.reshape(E, Np)
    face_ref = AqLe[:, ~interior].ravel()
    got = AqF.cpu().numpy()
    if layout == "SLOT":
        face_ref, got = face_ref[cols], got[:ncop]
    assert np.max(np.abs(got - face_ref)) <= 1e-12 * np.max(np.abs(AqL))
    rd, sd = dev(rows), dev(starts)
    cd = dev(cols) if layout == "ELEMENT" else None
    cptr = cd.data_ptr() if cd is not None else None
    scal = torch.zeros(16, dtype=torch.float64, device="cuda")
    scal[0] = 2.0  # r.z of the iteration: alpha = 2 / p.Ap
    lp().check(L.libp_pcg_gather_face_pap_f64(rows.size, rd.data_ptr(), sd.data_ptr(), cptr, AqF.data_ptr(),
                                              Aq.data_ptr(), qd.data_ptr(), nax, part.data_ptr(), scal.data_ptr(), st))
    torch.cuda.synchronize()
    y = Aq.cpu().numpy()
    assert np.all(np.isfinite(y))
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))
    pap = float(scal[4])
    assert abs(pap - q @ ref) <= 1e-12 * abs(q @ ref) * 10
    assert abs(float(scal[2]) - 2.0 / pap) <= 1e-15 * abs(2.0 / pap)
    # the plain face gather gives the same rows
    Aq2 = Aq.clone()
    lp().check(L.libp_ogs_gather_face_f64(rows.size, rd.data_ptr(), sd.data_ptr(), cptr, AqF.data_ptr(),
                                          Aq2.data_ptr(), st))
    torch.cuda.synchronize()
    assert torch.equal(Aq, Aq2)


@pytest.mark.parametrize("N", [3, 7])
@pytest.mark.parametrize("layout", ["SLOT", "ELEMENT"])
def test_split_pcg_matches_two_pass(N, layout):
    """Jacobi-PCG with the split operator and its fused p.Ap: the same iterates as the
    two-pass form up to the order of the p.Ap partial sums."""
    s = box(4, N, lam=0.0, curve=0.0, **{"LINEAR SOLVER": "PCG", "PRECONDITIONER": "JACOBI"})
    ra = lp().Elliptic(dict(s, **{"OPERATOR SPLIT GATHER": "TRUE", "OPERATOR FACE LAYOUT": layout})).run()
    rb = lp().Elliptic(dict(s, **{"OPERATOR SPLIT GATHER": "FALSE"})).run()
    print(f"N={N}: split {ra['iterations']} its {ra['norm']:.15f}; two-pass {rb['iterations']} its {rb['norm']:.15f}",
          flush=True)
    assert abs(ra["iterations"] - rb["iterations"]) <= 1
    assert abs(ra["norm"] - rb["norm"]) <= 1e-11 * abs(rb["norm"])

"""Row-block gathers on synthetic CSRs, bit for bit against sequential sums.

The gathers (ogs Trans gather, libs/ogs/okl/ogsKernels.okl:86-122, and the
kernels fused with it) stage each block of 512 rows through LDS, 2048 copies
at a time, and sum every row from zero in ascending column order. These CSRs
exercise blocks of one chunk and of several (rows that straddle chunks), row
counts around the block size, a single row, rows longer than eight copies, and
the fused forms (residual, add, Chebyshev, p.Ap, the Jacobi-PCG update),
against numpy sums taken in the same order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def lp():
    import libparanumal_amd
    return libparanumal_amd


def csr(Nrows, maxlen, rng, long_blocks=False):
    lens = rng.integers(1, maxlen + 1, size=Nrows)
    if long_blocks:  # some blocks of 512 rows with > 4096 copies (three chunks or more)
        for b in range(0, Nrows, 1024):
            lens[b:b + 512] = rng.integers(9, maxlen + 1, size=len(lens[b:b + 512]))
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    NL = int(rs[-1]) + 64
    # distinct ascending local ids per row, as an ogs CSR has: a random base plus a sorted
    # random subset of 64 offsets
    base = rng.integers(0, NL - 64, size=Nrows)
    perm = np.argsort(rng.random((Nrows, 64)), axis=1)
    take = np.arange(64)[None, :] < lens[:, None]
    chosen = np.sort(np.where(take, perm, 64), axis=1)  # each row's first lens offsets, ascending
    ci = (chosen + base[:, None])[np.arange(64)[None, :] < lens[:, None]].astype(np.int32)
    return rs, ci, NL, lens


def row_sums(rs, ci, v, lens):
    """Sequential sum of each row from zero in column order (cumsum is sequential)."""
    m = int(lens.max())
    pad = np.zeros((len(lens), m), v.dtype)
    for k in range(m):
        sel = lens > k
        pad[sel, k] = v[ci[rs[:-1][sel] + k]]
    return np.cumsum(pad, axis=1)[np.arange(len(lens)), lens - 1]


# (rows, longest row, blocks of 512 rows with more than two chunks of copies)
CASES = [(1, 8, False), (511, 8, False), (512, 8, False), (513, 8, False), (5000, 3, False), (40000, 8, False),
         (3001, 20, True)]


@pytest.mark.parametrize("Nrows,maxlen,long_blocks", CASES)
def test_plain_residual_add_gathers(Nrows, maxlen, long_blocks):
    rng = np.random.default_rng(Nrows + maxlen)
    rs, ci, NL, lens = csr(Nrows, maxlen, rng, long_blocks)
    if long_blocks:
        nb = (Nrows + 511) // 512
        assert max(rs[min((b + 1) * 512, Nrows)] - rs[b * 512] for b in range(nb)) > 4096
    L = lp().lib()
    st = torch.cuda.current_stream().cuda_stream
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rsd, cid = dev(rs), dev(ci)
    for dt in (np.float64, np.float32):
        v = rng.uniform(-1, 1, NL).astype(dt)
        r = rng.uniform(-1, 1, Nrows).astype(dt)
        g0 = rng.uniform(-1, 1, Nrows).astype(dt)
        ref = row_sums(rs, ci, v, lens)
        vd = dev(v)
        out = torch.full((Nrows,), float("nan"), dtype=vd.dtype, device="cuda")
        res = out.clone()
        add = dev(g0)
        f64 = dt == np.float64
        sfx = "f64" if f64 else "f32"
        lp().check(getattr(L, f"libp_ogs_gather_{sfx}")(Nrows, rsd.data_ptr(), cid.data_ptr(), vd.data_ptr(),
                                                       out.data_ptr(), st))
        lp().check(getattr(L, f"libp_ogs_gather_residual_{sfx}")(Nrows, rsd.data_ptr(), cid.data_ptr(),
                                                                vd.data_ptr(), dev(r).data_ptr(), res.data_ptr(),
                                                                st))
        lp().check(getattr(L, f"libp_ogs_gather_add_{sfx}")(Nrows, rsd.data_ptr(), cid.data_ptr(), vd.data_ptr(),
                                                           add.data_ptr(), st))
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref), (Nrows, sfx)
        assert np.array_equal(res.cpu().numpy(), (r - ref).astype(dt)), (Nrows, sfx)
        assert np.array_equal(add.cpu().numpy(), (ref + g0).astype(dt)), (Nrows, sfx)


# 1,200,000 rows: 2,344 row blocks, so the Jacobi update's partials are summed in two stages
@pytest.mark.parametrize("Nrows,maxlen,long_blocks", [(513, 8, False), (40000, 8, False), (3001, 20, True),
                                                      (1200000, 3, False)])
def test_fused_pcg_and_chebyshev_gathers(Nrows, maxlen, long_blocks):
    rng = np.random.default_rng(7 * Nrows)
    rs, ci, NL, lens = csr(Nrows, maxlen, rng, long_blocks)
    L = lp().lib()
    st = torch.cuda.current_stream().cuda_stream
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rsd, cid = dev(rs), dev(ci)
    v = rng.uniform(-1, 1, NL)
    p, x, r = rng.uniform(-1, 1, Nrows), rng.uniform(-1, 1, Nrows), rng.uniform(-1, 1, Nrows)
    invD = rng.uniform(0.5, 2, Nrows)
    ref = row_sums(rs, ci, v, lens)
    vd, pd, invd = dev(v), dev(p), dev(invD)
    # p.Ap gather (the four-launch PCG form): gather bit-exact, p.Ap to rounding
    gv = torch.full((Nrows,), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.zeros(16, dtype=torch.float64, device="cuda")
    s[0] = 3.0
    ws = torch.empty(L.libp_reduce_ws_bytes(), dtype=torch.uint8, device="cuda")
    lp().check(L.libp_pcg_gather_pap_f64(Nrows, rsd.data_ptr(), cid.data_ptr(), vd.data_ptr(), gv.data_ptr(),
                                         pd.data_ptr(), s.data_ptr(), ws.data_ptr(), st))
    torch.cuda.synchronize()
    assert np.array_equal(gv.cpu().numpy(), ref)
    pap = float(p @ ref)
    assert abs(float(s[4]) - pap) <= 1e-12 * abs(pap) * 10
    assert abs(float(s[2]) - 3.0 / float(s[4])) <= 1e-15 * abs(float(s[2]))
    # the gather fused with the Jacobi-PCG update, against the gather + update kernels
    part = torch.zeros(L.libp_pcg_gather_jacobi_partials(Nrows), dtype=torch.float64, device="cuda")
    xa, ra, za = dev(x), dev(r), torch.empty(Nrows, dtype=torch.float64, device="cuda")
    sa = torch.zeros(16, dtype=torch.float64, device="cuda")
    sa[0], sa[2] = 2.0, 0.3
    sb = sa.clone()
    lp().check(L.libp_pcg_gather_jacobi_update_f64(Nrows, rsd.data_ptr(), cid.data_ptr(), vd.data_ptr(),
                                                   invd.data_ptr(), pd.data_ptr(), xa.data_ptr(), ra.data_ptr(),
                                                   za.data_ptr(), sa.data_ptr(), part.data_ptr(), st))
    xb, rb, zb = dev(x), dev(r), torch.empty(Nrows, dtype=torch.float64, device="cuda")
    lp().check(L.libp_pcg_jacobi_update_f64(Nrows, invd.data_ptr(), pd.data_ptr(), gv.data_ptr(), xb.data_ptr(),
                                            rb.data_ptr(), zb.data_ptr(), sb.data_ptr(), ws.data_ptr(), st))
    torch.cuda.synchronize()
    assert torch.equal(xa, xb) and torch.equal(ra, rb) and torch.equal(za, zb)
    sah, sbh = sa.cpu().numpy(), sb.cpu().numpy()
    assert abs(sah[5] - sbh[5]) <= 1e-13 * sbh[5] and abs(sah[0] - sbh[0]) <= 1e-13 * abs(sbh[0])
    # Chebyshev start and step fused with the gather, against gather + the separate kernels
    for start in (True, False):
        RES, d, X = (dev(rng.uniform(-1, 1, Nrows)) for _ in range(3))
        RES2, d2, X2 = RES.clone(), d.clone(), X.clone()
        rv = dev(r)
        if start:
            lp().check(L.libp_ogs_gather_cheb_start_f64(Nrows, rsd.data_ptr(), cid.data_ptr(), vd.data_ptr(),
                                                        invd.data_ptr(), rv.data_ptr(), 0.7, RES.data_ptr(),
                                                        d.data_ptr(), X.data_ptr(), st))
            lp().check(L.libp_cheb_start_f64(Nrows, invd.data_ptr(), rv.data_ptr(), gv.data_ptr(), 0.7,
                                             RES2.data_ptr(), d2.data_ptr(), X2.data_ptr(), 0, st))
        else:
            lp().check(L.libp_ogs_gather_cheb_step_f64(Nrows, rsd.data_ptr(), cid.data_ptr(), vd.data_ptr(),
                                                       invd.data_ptr(), 0.4, 0.2, RES.data_ptr(), d.data_ptr(),
                                                       X.data_ptr(), st))
            lp().check(L.libp_cheb_step_f64(Nrows, invd.data_ptr(), gv.data_ptr(), 0.4, 0.2, RES2.data_ptr(),
                                            d2.data_ptr(), X2.data_ptr(), st))
        torch.cuda.synchronize()
        assert torch.equal(RES, RES2) and torch.equal(d, d2) and torch.equal(X, X2), start

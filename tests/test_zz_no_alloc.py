"""The allocation-free fused Jacobi-PCG update against the workspace form.

libp_pcg_jacobi_update_noalloc_f64 keeps its block partials in a static
device array instead of a caller workspace. Through the C ABI it must give
x, r, z and every scalar of libp_pcg_jacobi_update_f64 bit for bit, on two
calls in a row with the second on a smaller vector (the array keeps the first
call's partials), on the 16-byte and the unaligned paths."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(L, lp, N, seed, offset):
    g = torch.Generator(device="cuda").manual_seed(seed)

    def vec(lo, hi):
        # offset 1 gives 8-byte-aligned (not 16-byte-aligned) pointers: the scalar path
        base = torch.rand(N + offset, dtype=torch.float64, device="cuda", generator=g) * (hi - lo) + lo
        return base[offset:]

    invDiag, p, Ap, x, r = vec(0.5, 2.0), vec(-1, 1), vec(-1, 1), vec(-1, 1), vec(-1, 1)
    s0 = torch.zeros(16, dtype=torch.float64, device="cuda")
    s0[0] = 0.75   # r.z of the previous iteration
    s0[2] = 0.37   # alpha
    s0[9] = 4.0    # ITERS
    outs = []
    stream = torch.cuda.current_stream().cuda_stream
    for form in ("workspace", "noalloc"):
        xx, rr = x.clone(), r.clone()
        zz = torch.full_like(x, float("nan"))
        s = s0.clone()
        if form == "workspace":
            ws = torch.empty(L.libp_reduce_ws_bytes(), dtype=torch.uint8, device="cuda")
            st = L.libp_pcg_jacobi_update_f64(N, invDiag.data_ptr(), p.data_ptr(), Ap.data_ptr(), xx.data_ptr(),
                                              rr.data_ptr(), zz.data_ptr(), s.data_ptr(), ws.data_ptr(), stream)
        else:
            st = L.libp_pcg_jacobi_update_noalloc_f64(N, invDiag.data_ptr(), p.data_ptr(), Ap.data_ptr(),
                                                      xx.data_ptr(), rr.data_ptr(), zz.data_ptr(), s.data_ptr(),
                                                      stream)
        lp.check(st)
        torch.cuda.synchronize()
        outs.append((xx.cpu().numpy(), rr.cpu().numpy(), zz.cpu().numpy(), s.cpu().numpy()))
    (xa, ra, za, sa), (xb, rb, zb, sb) = outs
    print(f"N={N} offset={offset}: r.r {sb[5]:.17e} r.z {sb[0]:.17e}", flush=True)
    assert np.array_equal(xa, xb) and np.array_equal(ra, rb) and np.array_equal(za, zb)
    assert np.array_equal(sa, sb)
    # what the update computes, in numpy (dots to rounding)
    xh, rh = x.cpu().numpy(), r.cpu().numpy()
    r_new = rh - 0.37 * Ap.cpu().numpy()
    z_new = invDiag.cpu().numpy() * r_new
    assert np.allclose(xb, xh + 0.37 * p.cpu().numpy(), rtol=0, atol=1e-15)
    assert np.allclose(rb, r_new, rtol=0, atol=1e-15)
    assert abs(sb[5] - r_new @ r_new) <= 1e-12 * (r_new @ r_new)
    assert abs(sb[0] - r_new @ z_new) <= 1e-12 * abs(r_new @ z_new)
    assert sb[1] == 0.75 and sb[9] == 5.0 and sb[8] == 0.0
    assert abs(sb[3] - sb[0] / 0.75) <= 1e-15 * abs(sb[3])


def test_noalloc_jacobi_update_matches_workspace_form():
    import libparanumal_amd as lp
    L = lp.lib()
    assert L.libp_pcg_jacobi_update_noalloc_max_n() == 2**31 - 1
    # first a large vector (1024 partial blocks), then a smaller one (5 blocks), then an unaligned one
    _case(L, lp, 3_000_001, seed=1, offset=0)
    _case(L, lp, 4_097, seed=2, offset=0)
    _case(L, lp, 777_777, seed=3, offset=1)


def test_noalloc_rejects_bad_arguments():
    import libparanumal_amd as lp
    L = lp.lib()
    v = torch.zeros(8, dtype=torch.float64, device="cuda")
    ptr = v.data_ptr()
    assert L.libp_pcg_jacobi_update_noalloc_f64(-1, ptr, ptr, ptr, ptr, ptr, ptr, ptr, None) == 1
    assert b"N outside" in L.libp_last_error()
    assert L.libp_pcg_jacobi_update_noalloc_f64(8, None, ptr, ptr, ptr, ptr, ptr, ptr, None) == 1
    assert b"null" in L.libp_last_error()
    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()

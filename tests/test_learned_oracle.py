"""CPU checks of the learned smoother's oracle (no GPU).

The learned smoother has no reference counterpart, so its oracle is pinned by
an independent construction of the neighbour traces (box structure here,
face connectivity in the oracle), by finite differences of the loss (which pin
the whole backward pass), and by special cases with closed forms.
"""
import numpy as np
import pytest

from tests.oracle_lib import OracleElliptic, OracleLearned


def box(n, N, flag=1, curve=0.0, lam=0.0):
    return {"FORMAT": "2.0", "DATA FILE": "data/ellipticSine3D.h", "MESH FILE": "BOX", "MESH DIMENSION": 3,
            "ELEMENT TYPE": 12, "BOX NX": n, "BOX NY": n, "BOX NZ": n, "BOX DIMX": 2, "BOX DIMY": 2, "BOX DIMZ": 2,
            "BOX BOUNDARY FLAG": flag, "POLYNOMIAL DEGREE": N, "THREAD MODEL": "Serial", "PLATFORM NUMBER": 0,
            "DEVICE NUMBER": 0, "DISCRETIZATION": "CONTINUOUS", "LINEAR SOLVER": "PCG", "PRECONDITIONER": "NONE",
            "MULTIGRID SMOOTHER": "CHEBYSHEV", "PARALMOND CYCLE": "VCYCLE", "PARALMOND STRENGTH": "SYMMETRIC",
            "PARALMOND AGGREGATION": "UNSMOOTHED", "PARALMOND SMOOTHER": "CHEBYSHEV", "OUTPUT TO FILE": "FALSE",
            "VERBOSE": "FALSE", "LAMBDA": lam, "BOX CURVED MAP AMPLITUDE": curve}


def node(i, j, k, Nq):
    return i + Nq * (j + Nq * k)


def expected_trace(n, N, g2l):
    """Traces from the box structure: e = ex + n (ey + n ez); face f's nodes in
    increasing local index; the neighbour layer next to the shared face."""
    Nq, Np, Nfp = N + 1, (N + 1) ** 3, (N + 1) ** 2
    out = -np.ones((n ** 3, 6, Nfp), dtype=np.int64)
    step = {0: (0, 0, -1), 1: (0, -1, 0), 2: (1, 0, 0), 3: (0, 1, 0), 4: (-1, 0, 0), 5: (0, 0, 1)}
    for ez in range(n):
        for ey in range(n):
            for ex in range(n):
                e = ex + n * (ey + n * ez)
                for f in range(6):
                    dx, dy, dz = step[f]
                    qx, qy, qz = ex + dx, ey + dy, ez + dz
                    if not (0 <= qx < n and 0 <= qy < n and 0 <= qz < n):
                        continue
                    en = qx + n * (qy + n * qz)
                    for b in range(Nq):
                        for a in range(Nq):
                            if f == 0: inner = node(a, b, N - 1, Nq)
                            elif f == 1: inner = node(a, N - 1, b, Nq)
                            elif f == 2: inner = node(1, a, b, Nq)
                            elif f == 3: inner = node(a, 1, b, Nq)
                            elif f == 4: inner = node(N - 1, a, b, Nq)
                            else: inner = node(a, b, 1, Nq)
                            out[e, f, a + Nq * b] = g2l[en * Np + inner]
    return out.reshape(-1)


@pytest.mark.parametrize("n,N,curve", [(3, 2, 0.0), (4, 3, 0.05)])
def test_trace_map_matches_box_structure(n, N, curve):
    el = OracleElliptic(box(n, N, curve=curve))
    ol = OracleLearned(el, 4)
    g2l = el.array("GlobalToLocal")
    got = ol.trace()
    assert np.array_equal(got, expected_trace(n, N, g2l))
    # an interior element has all six neighbours, every trace node is a dof
    Nfp = (N + 1) ** 2
    e = 1 + n * (1 + n * 1)
    assert (got.reshape(-1, 6 * Nfp)[e] >= 0).all()


def random_case(n=3, N=2, H=6, seed=0, curve=0.05):
    el = OracleElliptic(box(n, N, curve=curve))
    ol = OracleLearned(el, H)
    rng = np.random.default_rng(seed)
    d = ol.dims
    th = np.zeros(d["n"])
    th[d["W1"]:d["b1"]] = rng.uniform(-1, 1, d["H"] * d["nIn"]) * np.sqrt(6 / d["nIn"])
    th[d["b1"]:d["W2"]] = rng.uniform(-0.1, 0.1, H)
    th[d["W2"]:d["b2"]] = rng.uniform(-1, 1, H * H) * np.sqrt(6 / H)
    th[d["b2"]:d["W3"]] = rng.uniform(-0.1, 0.1, H)
    th[d["W3"]:d["b3"]] = rng.uniform(-0.3, 0.3, d["nOut"] * H)
    th[d["b3"]:d["omega"]] = rng.uniform(-0.1, 0.1, d["nOut"])
    th[d["omega"]] = 0.4
    Ng = ol.Ng
    r = rng.standard_normal(Ng)
    eh = rng.standard_normal(Ng)
    invD = 1.0 / el.diagonal()
    nrm = float(eh @ el.operator(eh))
    return el, ol, th, r, eh, invD, nrm


def test_gradient_matches_finite_differences():
    el, ol, th, r, eh, invD, nrm = random_case()
    J, g = ol.loss_grad(th, r, eh, nrm, invD)
    assert J > 0
    d = ol.dims
    rng = np.random.default_rng(1)
    h = 1e-6
    for blk, (a, b) in {"W1": (d["W1"], d["b1"]), "b1": (d["b1"], d["W2"]), "W2": (d["W2"], d["b2"]),
                        "b2": (d["b2"], d["W3"]), "W3": (d["W3"], d["b3"]), "b3": (d["b3"], d["omega"]),
                        "omega": (d["omega"], d["n"])}.items():
        v = np.zeros(d["n"])
        v[a:b] = rng.standard_normal(b - a)
        Jp, _ = ol.loss_grad(th + h * v, r, eh, nrm, invD, grad=False)
        Jm, _ = ol.loss_grad(th - h * v, r, eh, nrm, invD, grad=False)
        fd = (Jp - Jm) / (2 * h)
        an = float(g @ v)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (blk, fd, an)


def test_planted_sign_error_is_caught():
    # the finite-difference check fails on a gradient with a flipped block
    el, ol, th, r, eh, invD, nrm = random_case(seed=3)
    _, g = ol.loss_grad(th, r, eh, nrm, invD)
    d = ol.dims
    v = np.zeros(d["n"])
    v[d["W3"]:d["b3"]] = np.random.default_rng(2).standard_normal(d["b3"] - d["W3"])
    h = 1e-6
    fd = (ol.loss_grad(th + h * v, r, eh, nrm, invD, grad=False)[0] -
          ol.loss_grad(th - h * v, r, eh, nrm, invD, grad=False)[0]) / (2 * h)
    bad = -float(g @ v)
    assert abs(fd - bad) > 1e-3 * abs(fd)


def test_zero_last_layer_is_damped_jacobi():
    el, ol, th, r, eh, invD, nrm = random_case(seed=4)
    d = ol.dims
    th[d["W3"]:d["omega"]] = 0.0
    _, _, _, c = ol.forward(th, r, invD)
    assert np.allclose(c, th[d["omega"]] * invD * r, rtol=1e-14, atol=0)


def test_positive_homogeneity_and_zero_loss_at_target():
    el, ol, th, r, eh, invD, nrm = random_case(seed=5)
    _, _, _, c1 = ol.forward(th, r, invD)
    _, _, _, c2 = ol.forward(th, 2.5 * r, invD)
    assert np.allclose(c2, 2.5 * c1, rtol=1e-12, atol=1e-14 * np.abs(c1).max())
    J, _ = ol.loss_grad(th, r, c1, float(c1 @ el.operator(c1)), invD, grad=False)
    assert abs(J) < 1e-24


def test_bf16_rounding_known_values():
    from tests.oracle_lib import load
    L = load()
    assert L.orc_bf16r(1.0) == 1.0
    assert L.orc_bf16r(1.0 + 2.0 ** -8) == 1.0          # tie rounds to even (1.0)
    assert L.orc_bf16r(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie rounds up to even
    assert L.orc_bf16r(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
    assert L.orc_bf16r(-3.14159) == -3.140625


def test_linear_identity_layer_is_jacobi():
    """Linear architecture (H = 0) with W = [I | 0] (the element's own inputs) and
    omega: each copy returns s_e z/s_e = z, so c = (1 + omega) D^{-1} r."""
    o = OracleElliptic(box(3, 3, flag=1, curve=0.05))
    ol = OracleLearned(o, 0)
    d = ol.dims
    th = np.zeros(d["n"])
    W = th[:d["omega"]].reshape(d["nOut"], d["nIn"])
    W[:, :d["nOut"]] = np.eye(d["nOut"])
    th[d["omega"]] = 0.25
    invD = 1.0 / o.diagonal()
    r = np.random.default_rng(3).standard_normal(ol.Ng)
    _, _, _, c = ol.forward(th, r, invD)
    assert np.allclose(c, 1.25 * invD * r, rtol=1e-13, atol=0)


def test_linear_gram_recovers_the_identity_map():
    """Least squares with targets eh = D^{-1} r: the element target eh_e / s_e is
    exactly the element's own block of its input, so the normal equations
    G W^T = B^T built from the oracle's Gram matrices return W = [I | 0]. A
    transposed or mis-indexed Gram matrix fails this."""
    o = OracleElliptic(box(3, 3, flag=1, curve=0.05))
    ol = OracleLearned(o, 0)
    d = ol.dims
    invD = 1.0 / o.diagonal()
    rng = np.random.default_rng(4)
    G = np.zeros((d["nIn"], d["nIn"]))
    B = np.zeros((d["nOut"], d["nIn"]))
    for _ in range(12):
        r = rng.standard_normal(ol.Ng)
        g, b = ol.linear_gram(r, invD * r, invD, bf16=False)
        G += g
        B += b
    assert np.allclose(G, G.T, rtol=0, atol=1e-12 * np.abs(G).max())
    W = np.linalg.solve(G + 1e-12 * np.trace(G) / d["nIn"] * np.eye(d["nIn"]), B.T).T
    ref = np.zeros_like(W)
    ref[:, :d["nOut"]] = np.eye(d["nOut"])
    # masked nodes contribute zero rows and columns; compare on the live inputs
    live = np.diag(G) > 0
    err = np.abs(W[:, live] - ref[:, live]).max()
    print("identity recovery error", err)
    assert err < 1e-6

"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline load this. The oracle restates the reference's Serial-mode path; see
oracle/oracle.hpp.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

c_int, c_long, c_double, c_void_p, c_char_p = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
_lib = None
_DT = {0: np.float64, 1: np.int32, 2: np.int64}


def build():
    subprocess.run(["make", "-C", ORACLE_DIR, "-s"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        sig = {
            "orc_last_error": (c_char_p, []),
            "orc_create": (c_void_p, [c_char_p]),
            "orc_free": (None, [c_void_p]),
            "orc_geti": (c_long, [c_void_p, c_char_p]),
            "orc_getd": (c_double, [c_void_p, c_char_p]),
            "orc_array": (c_long, [c_void_p, c_char_p, c_void_p, ctypes.POINTER(c_int)]),
            "orc_partial_ax": (None, [c_void_p, c_void_p, c_void_p]),
            "orc_operator": (None, [c_void_p, c_void_p, c_void_p]),
            "orc_gather": (None, [c_void_p, c_void_p, c_void_p, c_int]),
            "orc_scatter": (None, [c_void_p, c_void_p, c_void_p]),
            "orc_diagonal": (None, [c_void_p, c_void_p]),
            "orc_rhs_local": (None, [c_void_p, c_void_p]),
            "orc_inner": (c_double, [c_long, c_void_p, c_void_p]),
            "orc_pcg": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_int, c_int, c_int,
                                ctypes.POINTER(c_double)]),
            "orc_mg_setup": (c_int, [c_void_p]),
            "orc_mg_apply": (None, [c_void_p, c_void_p, c_void_p]),
            "orc_mg_level_d": (c_double, [c_void_p, c_int, c_char_p]),
            "orc_mg_level_ndofs": (c_long, [c_void_p, c_int]),
            "orc_mg_level_op": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int]),
            "orc_max_eig": (c_double, [c_void_p, c_long]),
            "orc_assembled_csr": (c_long, [c_void_p, c_void_p, c_void_p, c_void_p]),
            "orc_amg_setup": (c_int, [c_void_p, c_long]),
            "orc_amg_levels": (c_int, [c_void_p, c_int]),
            "orc_amg_level_d": (c_double, [c_void_p, c_int, c_int, c_char_p]),
            "orc_amg_level_array": (c_long, [c_void_p, c_int, c_int, c_char_p, c_void_p, ctypes.POINTER(c_int)]),
            "orc_amg_vcycle": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
            "orc_amg_smooth": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int]),
            "orc_run": (c_double, [c_char_p, ctypes.POINTER(c_int), ctypes.POINTER(c_long)]),
            "orc_gll": (None, [c_int, c_void_p, c_void_p]),
            "orc_dmatrix": (None, [c_int, c_void_p]),
            "orc_interp": (None, [c_int, c_int, c_void_p]),
            "orc_rand_seq": (None, [c_int, c_void_p]),
            "orc_drand48_seq": (None, [c_long, c_int, c_void_p]),
            "orc_hess_eigs": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
            "orc_time_partial_ax": (c_double, [c_void_p, c_int, c_int]),
            "orc_time_leg": (c_double, [c_void_p, c_int, c_int, c_int, ctypes.POINTER(c_int)]),
            "orc_partial_ax_omp": (None, [c_void_p, c_void_p, c_void_p]),
            "orc_omp_threads": (c_int, []),
            "orc_learned_trace": (c_long, [c_void_p, c_void_p]),
            "orc_learned_loss_grad": (c_double, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                                                 c_int, c_void_p, c_void_p]),
            "orc_learned_forward": (None, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                           c_void_p, c_void_p]),
            "orc_learned_linear_gram": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
            "orc_bf16r": (c_double, [c_double]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def rc(settings):
    return "".join(f"[{k}]\n{v}\n\n" for k, v in settings.items())


class OracleElliptic:
    def __init__(self, settings):
        self.L = load()
        text = settings if isinstance(settings, str) else rc(settings)
        self.h = self.L.orc_create(text.encode())
        if not self.h:
            raise RuntimeError("oracle create failed: " + self.L.orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_free(self.h)
            self.h = None

    def geti(self, name):
        return int(self.L.orc_geti(self.h, name.encode()))

    def array(self, name):
        dt = c_int()
        n = self.L.orc_array(self.h, name.encode(), None, ctypes.byref(dt))
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=_DT[dt.value])
        self.L.orc_array(self.h, name.encode(), out.ctypes.data, ctypes.byref(dt))
        return out

    def partial_ax(self, q):
        NL = self.geti("Nelements") * self.geti("Np")
        out = np.empty(NL)
        q = np.ascontiguousarray(q, dtype=np.float64)
        self.L.orc_partial_ax(self.h, q.ctypes.data, out.ctypes.data)
        return out

    def partial_ax_omp(self, q):
        NL = self.geti("Nelements") * self.geti("Np")
        out = np.empty(NL)
        q = np.ascontiguousarray(q, dtype=np.float64)
        self.L.orc_partial_ax_omp(self.h, q.ctypes.data, out.ctypes.data)
        return out

    def operator(self, q):
        out = np.empty(self.geti("Ngather"))
        q = np.ascontiguousarray(q, dtype=np.float64)
        self.L.orc_operator(self.h, q.ctypes.data, out.ctypes.data)
        return out

    def gather(self, vL, trans=True):
        out = np.empty(self.geti("Ngather"))
        vL = np.ascontiguousarray(vL, dtype=np.float64)
        self.L.orc_gather(self.h, vL.ctypes.data, out.ctypes.data, 1 if trans else 0)
        return out

    def scatter(self, g):
        NL = self.geti("Nelements") * self.geti("Np")
        out = np.zeros(NL)
        g = np.ascontiguousarray(g, dtype=np.float64)
        self.L.orc_scatter(self.h, g.ctypes.data, out.ctypes.data)
        return out

    def diagonal(self):
        out = np.empty(self.geti("Ngather"))
        self.L.orc_diagonal(self.h, out.ctypes.data)
        return out

    def rhs_local(self):
        NL = self.geti("Nelements") * self.geti("Np")
        out = np.empty(NL)
        self.L.orc_rhs_local(self.h, out.ctypes.data)
        return out

    def pcg(self, x, r, tol, maxit, jacobi=True, pure_relative=False):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        r = np.ascontiguousarray(r, dtype=np.float64).copy()
        rd = c_double()
        it = self.L.orc_pcg(self.h, x.ctypes.data, r.ctypes.data, tol, maxit, int(jacobi), int(pure_relative),
                            ctypes.byref(rd))
        return it, x, r, rd.value

    def mg_setup(self):
        n = self.L.orc_mg_setup(self.h)
        if n < 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return n

    def mg_apply(self, r):
        z = np.zeros(self.geti("Ngather"))
        r = np.ascontiguousarray(r, dtype=np.float64)
        self.L.orc_mg_apply(self.h, r.ctypes.data, z.ctypes.data)
        return z

    def mg_level_ndofs(self, l):
        return int(self.L.orc_mg_level_ndofs(self.h, l))

    def mg_level_d(self, l, name):
        return float(self.L.orc_mg_level_d(self.h, l, name.encode()))

    # operations of p-MG level l (orc_mg_level_op)
    MG_OPERATOR, MG_COARSEN, MG_PROLONGATE, MG_SMOOTH, MG_INVDIAG = range(5)

    def mg_level_op(self, l, op, vin, out=None, flag=0):
        n_out = {0: l, 1: l + 1, 2: l, 3: l, 4: l}[op]
        if out is None:
            out = np.zeros(self.mg_level_ndofs(n_out))
        out = np.ascontiguousarray(out, dtype=np.float64)
        vin = np.ascontiguousarray(vin if vin is not None else np.zeros(1), dtype=np.float64)
        if self.L.orc_mg_level_op(self.h, l, op, vin.ctypes.data, out.ctypes.data, int(flag)) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return out

    def max_eig(self, seed=0):
        return float(self.L.orc_max_eig(self.h, seed))

    # parAlmond AMG: which = 0 the AMG on this handle's operator, 1 the p-MG's base level
    def assembled_csr(self):
        nnz = self.L.orc_assembled_csr(self.h, None, None, None)
        if nnz < 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        rs = np.empty(self.geti("Ngather") + 1, dtype=np.int32)
        ci = np.empty(nnz, dtype=np.int32)
        v = np.empty(nnz)
        self.L.orc_assembled_csr(self.h, rs.ctypes.data, ci.ctypes.data, v.ctypes.data)
        return rs, ci, v

    def amg_setup(self, seed=0):
        n = self.L.orc_amg_setup(self.h, seed)
        if n < 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return n

    def amg_levels(self, which=0):
        return int(self.L.orc_amg_levels(self.h, which))

    def amg_level_d(self, l, name, which=0):
        return float(self.L.orc_amg_level_d(self.h, which, l, name.encode()))

    def amg_level_array(self, l, name, which=0):
        dt = c_int()
        n = self.L.orc_amg_level_array(self.h, which, l, name.encode(), None, ctypes.byref(dt))
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=_DT[dt.value])
        self.L.orc_amg_level_array(self.h, which, l, name.encode(), out.ctypes.data, ctypes.byref(dt))
        return out

    def amg_matrix(self, l, which="A", amg=0):
        """scipy CSR of level l's A, P or R."""
        import scipy.sparse as sp
        rs = self.amg_level_array(l, which + "_rs", amg)
        ci = self.amg_level_array(l, which + "_ci", amg)
        v = self.amg_level_array(l, which + "_v", amg)
        n = rs.size - 1
        m = (ci.max() + 1) if ci.size else 0
        if which == "A":
            m = n
        return sp.csr_matrix((v, ci, rs), shape=(n, m))

    def amg_vcycle(self, b, which=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        if self.L.orc_amg_vcycle(self.h, which, b.ctypes.data, x.ctypes.data) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return x

    def amg_smooth(self, l, b, x, x_is_zero, which=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        if self.L.orc_amg_smooth(self.h, which, l, b.ctypes.data, x.ctypes.data, int(x_is_zero)) != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return x


def learned_dims(Np, Nfp, H):
    """Offsets of the learned smoother's parameter vector (W1, b1, W2, b2, W3, b3, omega);
    H = 0 is the linear architecture (W3 [Np][nIn], omega)."""
    nIn, nOut = Np + 6 * Nfp, Np
    if H == 0:
        o = {"W1": 0, "b1": 0, "W2": 0, "b2": 0, "W3": 0, "b3": 0, "omega": nOut * nIn}
        o["n"] = o["omega"] + 1
        o["nIn"], o["nOut"], o["H"] = nIn, nOut, H
        return o
    o = {"W1": 0}
    o["b1"] = o["W1"] + H * nIn
    o["W2"] = o["b1"] + H
    o["b2"] = o["W2"] + H * H
    o["W3"] = o["b2"] + H
    o["b3"] = o["W3"] + nOut * H
    o["omega"] = o["b3"] + nOut
    o["n"] = o["omega"] + 1
    o["nIn"], o["nOut"], o["H"] = nIn, nOut, H
    return o


class OracleLearned:
    """The learned smoother's CPU restatement on an oracle elliptic handle."""

    def __init__(self, el, H):
        self.el, self.L, self.H = el, el.L, H
        self.Np, self.E, self.Ng = el.geti("Np"), el.geti("Nelements"), el.geti("Ngather")
        self.Nfp = round(self.Np ** (2.0 / 3.0))
        self.dims = learned_dims(self.Np, self.Nfp, H)

    def trace(self):
        n = self.L.orc_learned_trace(self.el.h, None)
        out = np.empty(n, dtype=np.int32)
        self.L.orc_learned_trace(self.el.h, out.ctypes.data)
        return out

    def forward(self, theta, r, invD, bf16=False):
        d = self.dims
        th = np.ascontiguousarray(theta, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        invD = np.ascontiguousarray(invD, dtype=np.float64)
        X = np.empty(self.E * d["nIn"]); s = np.empty(self.E); Y = np.empty(self.E * d["nOut"]); c = np.empty(self.Ng)
        self.L.orc_learned_forward(self.el.h, self.H, th.ctypes.data, r.ctypes.data, invD.ctypes.data, int(bf16),
                                   X.ctypes.data, s.ctypes.data, Y.ctypes.data, c.ctypes.data)
        return X.reshape(self.E, -1), s, Y.reshape(self.E, -1), c

    def linear_gram(self, r, eh, invD, bf16=True):
        """G = X^T X [nIn][nIn] and B = T^T X [Np][nIn] of one least-squares sample."""
        d = self.dims
        r = np.ascontiguousarray(r, dtype=np.float64)
        eh = np.ascontiguousarray(eh, dtype=np.float64)
        invD = np.ascontiguousarray(invD, dtype=np.float64)
        G = np.empty((d["nIn"], d["nIn"])); B = np.empty((d["nOut"], d["nIn"]))
        self.L.orc_learned_linear_gram(self.el.h, r.ctypes.data, eh.ctypes.data, invD.ctypes.data, int(bf16),
                                       G.ctypes.data, B.ctypes.data)
        return G, B

    def loss_grad(self, theta, r, eh, nrm, invD, bf16=False, grad=True):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        eh = np.ascontiguousarray(eh, dtype=np.float64)
        invD = np.ascontiguousarray(invD, dtype=np.float64)
        g = np.empty(self.dims["n"]) if grad else None
        J = self.L.orc_learned_loss_grad(self.el.h, self.H, th.ctypes.data, r.ctypes.data, eh.ctypes.data, nrm,
                                         invD.ctypes.data, int(bf16), g.ctypes.data if grad else None, None)
        return J, g


def run(settings):
    L = load()
    it, ng = c_int(), c_long()
    text = settings if isinstance(settings, str) else rc(settings)
    norm = L.orc_run(text.encode(), ctypes.byref(it), ctypes.byref(ng))
    if norm != norm:
        raise RuntimeError(L.orc_last_error().decode())
    return norm, it.value, ng.value

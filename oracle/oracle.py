"""ctypes binding of the CPU oracle (oracle/oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.  See oracle.c for the restated reference
functions and how each is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")

_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_dp = ctypes.POINTER(ctypes.c_double)
_c_fp = ctypes.POINTER(ctypes.c_float)
_c_ip = ctypes.POINTER(ctypes.c_int)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.or_new.restype = ctypes.c_void_p
        L.or_free.argtypes = [ctypes.c_void_p]
        L.or_set_nodes.argtypes = [ctypes.c_void_p, ctypes.c_int64, _c_dp]
        L.or_add_block.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _c_i64p]
        L.or_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_ip, _c_ip]
        L.or_add_dirichlet.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_ip, _c_ip, ctypes.c_double, ctypes.c_double]
        L.or_add_bilinear.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_ip, _c_ip, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, _c_ip, ctypes.c_int, _c_dp, ctypes.c_double, ctypes.c_double]
        L.or_add_linear.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_ip, _c_ip, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _c_ip, ctypes.c_int,
                                    _c_dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.or_set_options.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.or_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_pre.argtypes = [ctypes.c_void_p]
        L.or_n_unknown.argtypes = [ctypes.c_void_p]
        L.or_n_unknown.restype = ctypes.c_int64
        L.or_n_fixed.argtypes = [ctypes.c_void_p]
        L.or_n_fixed.restype = ctypes.c_int64
        L.or_nnz.argtypes = [ctypes.c_void_p]
        L.or_nnz.restype = ctypes.c_int64
        L.or_get_csr.argtypes = [ctypes.c_void_p, _c_i64p, _c_i64p]
        L.or_get_block_dofs.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_i64p, _c_i32p]
        L.or_get_node_dof.argtypes = [ctypes.c_void_p, ctypes.c_int64, _c_i64p, _c_i32p]
        L.or_assemble.argtypes = [ctypes.c_void_p]
        L.or_get_values.argtypes = [ctypes.c_void_p, _c_dp, _c_dp]
        L.or_n_rhs.argtypes = [ctypes.c_void_p]
        L.or_get_rhs.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_dp, _c_dp]
        L.or_hilbert_key.argtypes = [ctypes.c_int, _c_fp, _c_fp, _c_fp]
        L.or_hilbert_key.restype = ctypes.c_uint64
        L.or_gauss_rule.argtypes = [ctypes.c_int, ctypes.c_int, _c_dp, _c_dp, ctypes.c_int]
        L.or_lagrange.argtypes = [ctypes.c_int, ctypes.c_int, _c_dp, _c_dp, _c_dp]
        L.or_elem_nodes.argtypes = [ctypes.c_int, _c_ip, _c_ip]
        L.or_ref_nodes.argtypes = [ctypes.c_int, ctypes.c_int, _c_ip]
        L.or_hilbert_index.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
        L.or_hilbert_index.restype = ctypes.c_uint64
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _ints(v):
    return (ctypes.c_int * max(len(v), 1))(*v)


def _program(coef):
    """(re, im, nops, ops, nconst, consts) of a coefficient: a number or an
    object with .is_constant/.value/.ops/.consts (gfem.ScalarFunction)."""
    if isinstance(coef, (int, float, complex)):
        c = complex(coef)
        return c.real, c.imag, 0, None, 0, None
    if coef.is_constant:
        c = complex(coef.value)
        return c.real, c.imag, 0, None, 0, None
    ops = np.ascontiguousarray(coef.ops, dtype=np.int32)
    consts = np.ascontiguousarray(coef.consts, dtype=np.float64)
    return 0.0, 0.0, len(ops) // 2, ops, len(consts) // 2, consts


class Problem:
    """One oracle problem (mesh + field + constraints + terms)."""

    def __init__(self):
        self.L = lib()
        self.h = ctypes.c_void_p(self.L.or_new())
        self.blocks = []
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_free(self.h)
            self.h = None

    def set_nodes(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.L.or_set_nodes(self.h, xyz.shape[0], _p(xyz, _c_dp))

    def add_block(self, dim, tag, gmsh_type, conn):
        conn = np.ascontiguousarray(conn, dtype=np.int64)
        ne = conn.shape[0] if conn.ndim == 2 else 0
        b = self.L.or_add_block(self.h, dim, tag, gmsh_type, ne, _p(conn, _c_i64p))
        if b < 0:
            raise ValueError("or_add_block failed")
        self.blocks.append((dim, tag, gmsh_type, conn.shape))
        return b

    def set_field(self, order, ents):
        d, t = zip(*ents)
        self.L.or_set_field(self.h, order, len(ents), _ints(d), _ints(t))

    def add_dirichlet(self, ents, value):
        d, t = zip(*ents)
        v = complex(value)
        self.L.or_add_dirichlet(self.h, len(ents), _ints(d), _ints(t), v.real, v.imag)

    def add_bilinear(self, ents, gauss, form, coef=1.0, tf_coef=1.0, hermitian=True, dt=0):
        d, t = zip(*ents)
        re, im, nops, ops, nc, consts = _program(coef)
        self._keep += [ops, consts]
        tc = complex(tf_coef)
        r = self.L.or_add_bilinear(self.h, len(ents), _ints(d), _ints(t), gauss, form, form, int(hermitian), dt, re,
                                   im, nops, None if ops is None else _p(ops, _c_ip), nc,
                                   None if consts is None else _p(consts, _c_dp), tc.real, tc.imag)
        if r < 0:
            raise ValueError("or_add_bilinear failed")

    def add_linear(self, ents, gauss, source, tf_coef=1.0, hermitian=True, rhs_idx=0):
        d, t = zip(*ents)
        re, im, nops, ops, nc, consts = _program(source)
        self._keep += [ops, consts]
        tc = complex(tf_coef)
        r = self.L.or_add_linear(self.h, len(ents), _ints(d), _ints(t), gauss, 0, int(hermitian), re, im, nops,
                                 None if ops is None else _p(ops, _c_ip), nc,
                                 None if consts is None else _p(consts, _c_dp), tc.real, tc.imag, rhs_idx)
        if r < 0:
            raise ValueError("or_add_linear failed")

    def set_options(self, symmetric=False, frequency=0.0, dofs_sort="default"):
        self.L.or_set_options(self.h, int(symmetric), frequency, 0 if dofs_sort in ("default", "hilbert") else 1)

    def set_threads(self, n):
        """Element-loop threads of assemble(): 1 is the sequential checker; more
        threads scatter with atomic adds (unspecified order, timing only)."""
        self.L.or_set_threads(self.h, int(n))

    def pre(self):
        if self.L.or_pre(self.h) != 0:
            raise RuntimeError("or_pre failed")

    @property
    def n_unknown(self):
        return self.L.or_n_unknown(self.h)

    @property
    def n_fixed(self):
        return self.L.or_n_fixed(self.h)

    @property
    def nnz(self):
        return self.L.or_nnz(self.h)

    def csr(self):
        ai = np.zeros(self.n_unknown + 1, np.int64)
        aj = np.zeros(max(self.nnz, 1), np.int64)
        self.L.or_get_csr(self.h, _p(ai, _c_i64p), _p(aj, _c_i64p))
        return ai, aj[: self.nnz]

    def block_maps(self, index):
        _, _, _, shape = self.blocks[index]
        num = np.zeros(shape, np.int64)
        typ = np.zeros(shape, np.int32)
        if self.L.or_get_block_dofs(self.h, index, _p(num, _c_i64p), _p(typ, _c_i32p)) != 0:
            raise RuntimeError("block maps unavailable")
        return num, typ

    def node_dof(self, tag):
        n = ctypes.c_int64()
        t = ctypes.c_int32()
        self.L.or_get_node_dof(self.h, tag, ctypes.byref(n), ctypes.byref(t))
        return n.value, t.value

    def assemble(self):
        r = self.L.or_assemble(self.h)
        if r != 0:
            raise RuntimeError(f"or_assemble failed ({r})")

    def values(self):
        re = np.zeros(max(self.nnz, 1))
        im = np.zeros(max(self.nnz, 1))
        self.L.or_get_values(self.h, _p(re, _c_dp), _p(im, _c_dp))
        return (re + 1j * im)[: self.nnz]

    def rhs(self, idx=0):
        re = np.zeros(max(self.n_unknown, 1))
        im = np.zeros(max(self.n_unknown, 1))
        if self.L.or_get_rhs(self.h, idx, _p(re, _c_dp), _p(im, _c_dp)) != 0:
            raise RuntimeError("rhs unavailable")
        return (re + 1j * im)[: self.n_unknown]


def hilbert_key(dim, f, mn, mx):
    L = lib()
    a = (ctypes.c_float * 3)(*f)
    b = (ctypes.c_float * 3)(*mn)
    c = (ctypes.c_float * 3)(*mx)
    return L.or_hilbert_key(dim, a, b, c)


def hilbert_index(dim, u, top):
    L = lib()
    a = (ctypes.c_uint32 * 3)(*(list(u) + [0] * (3 - len(u))))
    return L.or_hilbert_index(dim, a, top)


def ref_nodes(dim, order):
    L = lib()
    out = (ctypes.c_int * (4 * 64))()
    n = L.or_ref_nodes(dim, order, out)
    return [list(out[4 * a: 4 * a + 4]) for a in range(n)]


def gauss_rule(dim, degree, maxq=4096):
    L = lib()
    pts = np.zeros(3 * maxq)
    w = np.zeros(maxq)
    q = L.or_gauss_rule(dim, degree, _p(pts, _c_dp), _p(w, _c_dp), maxq)
    if q <= 0:
        raise ValueError("no rule")
    return pts[: 3 * q].reshape(q, 3), w[:q]


def lagrange(dim, order, uvw):
    L = lib()
    u = np.ascontiguousarray(uvw, dtype=np.float64)
    val = np.zeros(64)
    grad = np.zeros(3 * 64)
    n = L.or_lagrange(dim, order, _p(u, _c_dp), _p(val, _c_dp), _p(grad, _c_dp))
    return val[:n], grad[: 3 * n].reshape(n, 3)


def from_spec(spec):
    """Oracle problem of a workload spec (gfem.workloads.Spec)."""
    P = Problem()
    m = spec.model
    P.set_nodes(m.xyz())
    for b in m.blocks():
        P.add_block(b["dim"], b["tag"], b["type"], b["conn"])
    P.set_field(spec.order, spec.field_domain)
    for ents, value in spec.dirichlet:
        P.add_dirichlet(ents, value)
    for t in spec.terms:
        if t["kind"] == "bilinear":
            P.add_bilinear(t["domain"], t["gauss"], t["form"], t.get("coef", 1.0), t.get("tf_coef", 1.0),
                           t.get("hermitian", True), t.get("dt", 0))
        else:
            P.add_linear(t["domain"], t["gauss"], t["source"], t.get("tf_coef", 1.0), t.get("hermitian", True),
                         t.get("rhs_idx", 0))
    P.set_options(spec.symmetric, spec.frequency, spec.dofs_sort)
    return P

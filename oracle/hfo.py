"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module; the product (``higflow_amd``) never does.  The oracle is a
plain-C restatement of the reference algorithm (see ``oracle/hfo_internal.h``).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class Params(C.Structure):
    _fields_ = [
        ("Re", C.c_double), ("dt", C.c_double), ("t", C.c_double),
        ("projtype", C.c_int), ("flowtype", C.c_int), ("tempdiscrtype", C.c_int),
        ("convecdiscrtype", C.c_int), ("secondconvecdiscrtype", C.c_int),
        ("ve_model", C.c_int), ("ve_convec", C.c_int), ("kernel_transform", C.c_int),
        ("De", C.c_double), ("beta", C.c_double), ("kernel_tol", C.c_double),
        ("giesekus_alpha", C.c_double),
        ("rtol", C.c_double), ("atol", C.c_double), ("maxit", C.c_int),
        ("pin_pressure", C.c_double), ("order", C.c_int),
    ]


def params_from(case_params):
    p = Params()
    for name, _ in Params._fields_:
        if name in case_params:
            setattr(p, name, case_params[name])
    return p


_BP = C.CFUNCTYPE(C.c_double, C.c_int, C.POINTER(C.c_double), C.c_double)
_BV = C.CFUNCTYPE(C.c_double, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_double)
_SRC = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_double)
_FSRC = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_int, C.c_double)


class Functions(C.Structure):
    """hfo_functions (oracle/hfo.h): the higflow_functions callbacks."""
    _fields_ = [("get_boundary_pressure", _BP), ("get_boundary_velocity", _BV),
                ("get_source_term", _SRC), ("get_facet_source_term", _FSRC)]


def make_functions(dim, boundary_pressure=None, boundary_velocity=None, source=None, facet_source=None):
    pt = lambda c: np.ctypeslib.as_array(c, shape=(dim,)).copy()
    f = Functions()
    if boundary_pressure:
        f.get_boundary_pressure = _BP(lambda i, c, t: float(boundary_pressure(i, pt(c), t)))
    if boundary_velocity:
        f.get_boundary_velocity = _BV(lambda i, c, d, t: float(boundary_velocity(i, pt(c), d, t)))
    if source:
        f.get_source_term = _SRC(lambda c, t: float(source(pt(c), t)))
    if facet_source:
        f.get_facet_source_term = _FSRC(lambda c, d, t: float(facet_source(pt(c), d, t)))
    return f


_LIBS = {}


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib(dim, variant=""):
    key = (dim, variant)
    if key in _LIBS:
        return _LIBS[key]
    path = os.path.join(HERE, "liboracle%dd%s.so" % (dim, variant))
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    vp, dp, ip, cp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_char_p
    L.hfo_create.restype = vp
    L.hfo_destroy.argtypes = [vp]
    L.hfo_add_domain_amr.argtypes = [vp, cp]
    L.hfo_add_bc_amr.argtypes = [vp, C.c_int, cp, C.c_int, ip]
    L.hfo_setup.argtypes = [vp, C.POINTER(Params)]
    L.hfo_set_time.argtypes = [vp, C.c_double]
    L.hfo_get_time.argtypes = [vp]
    L.hfo_get_time.restype = C.c_double
    for f in ("hfo_field_size",):
        getattr(L, f).argtypes = [vp, C.c_int, C.c_int]
    for f in ("hfo_field_centers", "hfo_get_field", "hfo_set_field"):
        getattr(L, f).argtypes = [vp, C.c_int, C.c_int, dp]
    L.hfo_bc_count.argtypes = [vp, C.c_int, C.c_int]
    L.hfo_bc_centers.argtypes = [vp, C.c_int, C.c_int, dp]
    L.hfo_bc_set_values.argtypes = [vp, C.c_int, C.c_int, dp]
    L.hfo_set_functions.argtypes = [vp, vp]
    L.hfo_bc_set_valuetype.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.hfo_step.argtypes = [vp]
    L.hfo_step_viscoelastic.argtypes = [vp]
    L.hfo_advance.argtypes = [vp, C.c_int]
    L.hfo_stencil.argtypes = [vp, C.c_int, C.c_int, dp, dp, C.c_double, C.c_int, ip, dp, dp]
    L.hfo_facet_flag.argtypes = [vp, C.c_int, dp]
    L.hfo_wls_weights.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp]
    L.hfo_last_solve.argtypes = [vp, ip, dp]
    L.hfo_pinned_cell.argtypes = [vp]
    L.hfo_jacobi.argtypes = [dp, dp, dp]
    L.hfo_ve_local_rhs.argtypes = [C.POINTER(Params), dp, dp, dp]
    _LIBS[key] = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One oracle simulation built from a case dict (see higflow_amd.cases)."""

    def __init__(self, case, variant=""):
        self.dim = case["dim"]
        self.L = lib(self.dim, variant)
        self.h = self.L.hfo_create()
        self.case = case
        for amr in case["domains"]:
            if self.L.hfo_add_domain_amr(self.h, amr.encode()) < 0:
                raise RuntimeError("bad domain amr")
        for bc in case["bcs"]:
            ut = (C.c_int * self.dim)(*bc["utypes"])
            self.L.hfo_add_bc_amr(self.h, bc["id"], bc["amr"].encode(), bc["ptype"], ut)
        self.par = params_from(case["params"])
        rc = self.L.hfo_setup(self.h, C.byref(self.par))
        if rc != 0:
            raise RuntimeError("hfo_setup failed %d" % rc)
        for i, bc in enumerate(case["bcs"]):
            for kind in range(self.dim + 1):
                xyz = self.bc_centers(kind, i)
                vals = np.asarray([bc["value"](kind, x) for x in xyz], dtype=np.float64)
                self.L.hfo_bc_set_values(self.h, kind, i, _dp(np.ascontiguousarray(vals)))

    def close(self):
        if self.h:
            self.L.hfo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self, field, comp=0):
        return self.L.hfo_field_size(self.h, field, comp)

    def centers(self, field, comp=0):
        n = self.size(field, comp)
        out = np.zeros((n, self.dim))
        self.L.hfo_field_centers(self.h, field, comp, _dp(out))
        return out

    def get(self, field, comp=0):
        n = self.size(field, comp)
        out = np.zeros(n)
        self.L.hfo_get_field(self.h, field, comp, _dp(out))
        return out

    def set(self, field, comp, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.size(field, comp)
        self.L.hfo_set_field(self.h, field, comp, _dp(v))

    def bc_centers(self, kind, i):
        n = self.L.hfo_bc_count(self.h, kind, i)
        out = np.zeros((max(n, 0), self.dim))
        if n > 0:
            self.L.hfo_bc_centers(self.h, kind, i, _dp(out))
        return out

    def set_bc(self, kind, i, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == len(self.bc_centers(kind, i))
        self.L.hfo_bc_set_values(self.h, kind, i, _dp(v))

    def set_functions(self, **callables):
        """The same callbacks as higflow_amd.native.Flow.set_functions (the
        structs have one layout)."""
        self._functions = make_functions(self.dim, **callables)
        self.L.hfo_set_functions(self.h, C.addressof(self._functions))

    def set_valuetype(self, kind, i, timedependent):
        assert self.L.hfo_bc_set_valuetype(self.h, kind, i, int(timedependent)) == 0

    def step(self):
        if self.case["params"].get("flowtype", 0) == 3:
            return self.L.hfo_step_viscoelastic(self.h)
        return self.L.hfo_step(self.h)

    def advance(self, n):
        return self.L.hfo_advance(self.h, n)

    def stencil(self, domain, dim, center, x, alpha=1.0):
        ids = (C.c_int * 4096)()
        w = np.zeros(4096)
        rhs = C.c_double(0)
        c = np.ascontiguousarray(center, dtype=np.float64)
        xx = np.ascontiguousarray(x, dtype=np.float64)
        n = self.L.hfo_stencil(self.h, domain, dim, _dp(c), _dp(xx), alpha, 4096, ids, _dp(w),
                               C.byref(rhs))
        return np.array(ids[:n], dtype=np.int64), w[:n].copy(), rhs.value

    def facet_flag(self, dim, x):
        xx = np.ascontiguousarray(x, dtype=np.float64)
        return self.L.hfo_facet_flag(self.h, dim, _dp(xx))

    def last_solve(self):
        it = C.c_int(0)
        r = C.c_double(0)
        conv = self.L.hfo_last_solve(self.h, C.byref(it), C.byref(r))
        return conv, it.value, r.value

    def pinned_cell(self):
        return self.L.hfo_pinned_cell(self.h)

    def time(self):
        return self.L.hfo_get_time(self.h)


def wls_weights(dim, ord_, maxpts, pts, x, libdim=2):
    L = lib(libdim, "")
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    xx = np.zeros(max(libdim, dim))
    xx[: len(x)] = x
    w = np.zeros(len(pts))
    n = L.hfo_wls_weights(dim, ord_, maxpts, len(pts), _dp(pts), _dp(xx), _dp(w))
    if n < 0:
        raise ValueError("bad wls input")
    return w


def jacobi(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    d = A.shape[0]
    L = lib(d, "")
    lam = np.zeros(d)
    R = np.zeros((d, d))
    L.hfo_jacobi(_dp(A), _dp(lam), _dp(R))
    return lam, R


def ve_local_rhs(params, Du, K):
    Du = np.ascontiguousarray(Du, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.float64)
    d = Du.shape[0]
    L = lib(d, "")
    out = np.zeros((d, d))
    p = params_from(params)
    L.hfo_ve_local_rhs(C.byref(p), _dp(Du), _dp(K), _dp(out))
    return out

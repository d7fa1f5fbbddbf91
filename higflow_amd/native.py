"""Python host mirror of the HiG-Flow projection step and solver-plugin API,
bound to ``libhigflow_hip.so`` (the HIP/gfx950 library; C ABI in
``include/higflow_hip.h``).

There is no CPU fallback: every call goes through the HIP library, and
importing the binding raises when the library is missing or no GPU is
visible for a call that needs one.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libhigflow_hip.so")

STAGES = ("interp", "momentum", "divergence", "solve", "project", "viscoelastic")


class HipError(RuntimeError):
    pass


class Params(C.Structure):
    """hfh_params (include/higflow_hip.h)."""
    _fields_ = [
        ("Re", C.c_double), ("dt", C.c_double), ("t", C.c_double),
        ("projtype", C.c_int), ("flowtype", C.c_int), ("tempdiscrtype", C.c_int),
        ("convecdiscrtype", C.c_int), ("secondconvecdiscrtype", C.c_int),
        ("ve_model", C.c_int), ("ve_convec", C.c_int), ("kernel_transform", C.c_int),
        ("De", C.c_double), ("beta", C.c_double), ("kernel_tol", C.c_double),
        ("giesekus_alpha", C.c_double),
        ("rtol", C.c_double), ("atol", C.c_double), ("maxit", C.c_int),
        ("pin_pressure", C.c_double), ("order", C.c_int),
        ("precision", C.c_int), ("inner_rtol", C.c_double),
    ]


_LIB = None


BP_FN = C.CFUNCTYPE(C.c_double, C.c_int, C.POINTER(C.c_double), C.c_double)
BV_FN = C.CFUNCTYPE(C.c_double, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_double)
SRC_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_double)
FSRC_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_int, C.c_double)


class Functions(C.Structure):
    """hfh_functions: the higflow_functions callbacks a step re-evaluates."""
    _fields_ = [("get_boundary_pressure", BP_FN), ("get_boundary_velocity", BV_FN),
                ("get_source_term", SRC_FN), ("get_facet_source_term", FSRC_FN)]


def make_functions(dim, boundary_pressure=None, boundary_velocity=None, source=None, facet_source=None):
    """Wrap Python callables as hfh_functions: boundary_pressure(id, x, t),
    boundary_velocity(id, x, d, t), source(x, t), facet_source(x, d, t),
    with x a numpy copy of the point.  Keep the returned object alive while
    the solver may call it."""
    pt = lambda c: np.ctypeslib.as_array(c, shape=(dim,)).copy()
    f = Functions()
    if boundary_pressure:
        f.get_boundary_pressure = BP_FN(lambda i, c, t: float(boundary_pressure(i, pt(c), t)))
    if boundary_velocity:
        f.get_boundary_velocity = BV_FN(lambda i, c, d, t: float(boundary_velocity(i, pt(c), d, t)))
    if source:
        f.get_source_term = SRC_FN(lambda c, t: float(source(pt(c), t)))
    if facet_source:
        f.get_facet_source_term = FSRC_FN(lambda c, d, t: float(facet_source(pt(c), d, t)))
    return f


def load():
    """Load libhigflow_hip.so (raises HipError when it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise HipError("libhigflow_hip.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIBPATH)
    vp, dp, ip, cp, lp = (C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_char_p,
                          C.POINTER(C.c_long))
    L.hfh_last_error.restype = C.c_char_p
    L.hfh_last_error.argtypes = []
    L.hfh_slv_create.restype = vp
    L.hfh_slv_create.argtypes = [C.c_int, C.c_int]
    L.hfh_slv_set_maxnonzeros.argtypes = [vp, C.c_int]
    L.hfh_slv_set_Ai.argtypes = [vp, C.c_int, C.c_int, ip, dp]
    for f in ("hfh_slv_set_bi", "hfh_slv_add_bi", "hfh_slv_set_xi", "hfh_slv_impose_value"):
        getattr(L, f).argtypes = [vp, C.c_int, C.c_double]
    L.hfh_slv_set_x.argtypes = [vp, dp]
    L.hfh_slv_get_x.argtypes = [vp, dp]
    L.hfh_slv_get_xi.argtypes = [vp, C.c_int]
    L.hfh_slv_get_xi.restype = C.c_double
    L.hfh_slv_get_x_scatter.argtypes = [vp, C.c_int, ip, dp]
    for f in ("hfh_slv_assemble_matrix", "hfh_slv_assemble_rhs", "hfh_slv_assemble_output",
              "hfh_slv_solve", "hfh_slv_destroy", "hfh_slv_iterations", "hfh_slv_converged"):
        getattr(L, f).argtypes = [vp]
    L.hfh_slv_destroy.restype = None
    L.hfh_slv_residual.argtypes = [vp]
    L.hfh_slv_residual.restype = C.c_double
    L.hfh_slv_set_tolerances.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.hfh_create.restype = vp
    L.hfh_create.argtypes = [C.c_int]
    L.hfh_destroy.argtypes = [vp]
    L.hfh_destroy.restype = None
    L.hfh_add_domain_amr.argtypes = [vp, cp]
    L.hfh_add_bc_amr.argtypes = [vp, C.c_int, cp, C.c_int, ip]
    L.hfh_setup.argtypes = [vp, C.POINTER(Params)]
    L.hfh_bc_count.argtypes = [vp, C.c_int, C.c_int]
    L.hfh_bc_centers.argtypes = [vp, C.c_int, C.c_int, dp]
    L.hfh_bc_set_values.argtypes = [vp, C.c_int, C.c_int, dp]
    L.hfh_set_functions.argtypes = [vp, C.POINTER(Functions)]
    L.hfh_bc_set_valuetype.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.hfh_field_size.argtypes = [vp, C.c_int, C.c_int]
    for f in ("hfh_field_centers", "hfh_get_field", "hfh_set_field"):
        getattr(L, f).argtypes = [vp, C.c_int, C.c_int, dp]
    L.hfh_field_range.argtypes = [vp, C.c_int, C.c_int, dp, dp]
    L.hfh_range_noalloc.argtypes = [vp, C.c_long, vp, vp]
    L.hfh_cell_bounds.argtypes = [vp, dp]
    L.hfh_solver_step.argtypes = [vp]
    L.hfh_solver_step_viscoelastic.argtypes = [vp]
    L.hfh_set_time.argtypes = [vp, C.c_double]
    L.hfh_get_time.argtypes = [vp]
    L.hfh_get_time.restype = C.c_double
    L.hfh_pinned_cell.argtypes = [vp]
    L.hfh_last_solve.argtypes = [vp, ip, dp, ip]
    L.hfh_stencil.argtypes = [vp, C.c_int, C.c_int, dp, dp, C.c_double, C.c_int, ip, dp, dp]
    L.hfh_timing.argtypes = [vp, dp, lp]
    L.hfh_reset_timing.argtypes = [vp]
    L.hfh_stage_bytes.argtypes = [vp, C.c_int]
    L.hfh_stage_bytes.restype = C.c_long
    L.hfh_sync.argtypes = [vp]
    L.hfh_operator_csr.argtypes = [vp, C.c_int, ip, ip, ip, ip, ip, dp]
    L.hfh_tables_digest.argtypes = [vp, C.POINTER(C.c_ulonglong)]
    L.hfh_solve_stats.argtypes = [vp, C.c_int, ip, dp]
    L.hfh_set_timing.argtypes = [vp, C.c_int]
    L.hfh_probe.argtypes = [vp, C.c_int, C.c_int, dp, lp]
    L.hfh_version.argtypes = []
    L.hfh_device_count.argtypes = []
    L.hfh_create.argtypes = [C.c_int]
    L.hfh_create_host.restype = vp
    L.hfh_create_host.argtypes = [C.c_int]
    L.hfh_create_batch.restype = vp
    L.hfh_create_batch.argtypes = [C.c_int, C.POINTER(vp), C.c_int]
    L.hfh_batch_size.argtypes = [vp]
    L.hfh_batch_field_size.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.hfh_batch_get_field.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.hfh_batch_set_field.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.hfh_batch_bc_set_values.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.hfh_batch_solve_stats.argtypes = [vp, C.c_int, C.c_int, ip, dp]
    _LIB = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _check(L, rc, what):
    if rc < 0:
        raise HipError("%s failed (%d): %s" % (what, rc, L.hfh_last_error().decode()))
    return rc


def params_from(case_params, **over):
    p = Params()
    d = dict(case_params)
    d.update(over)
    for name, _ in Params._fields_:
        if name in d:
            setattr(p, name, d[name])
    return p


class Flow:
    """Device-resident HiG-Flow solver built from a case dict
    (``higflow_amd.cases``); mirrors higflow_solver_step(_viscoelastic)."""

    def __init__(self, case, host_only=False, **param_over):
        self.L = load()
        self.dim = case["dim"]
        self.case = case
        self.h = (self.L.hfh_create_host if host_only else self.L.hfh_create)(self.dim)
        if not self.h:
            raise HipError("hfh_create: " + self.L.hfh_last_error().decode())
        for amr in case["domains"]:
            _check(self.L, self.L.hfh_add_domain_amr(self.h, amr.encode()), "hfh_add_domain_amr")
        for bc in case["bcs"]:
            ut = (C.c_int * self.dim)(*bc["utypes"])
            _check(self.L, self.L.hfh_add_bc_amr(self.h, bc["id"], bc["amr"].encode(), bc["ptype"], ut),
                   "hfh_add_bc_amr")
        self.params = params_from(case["params"], **param_over)
        _check(self.L, self.L.hfh_setup(self.h, C.byref(self.params)), "hfh_setup")
        for i, bc in enumerate(case["bcs"]):
            for kind in range(self.dim + 1):
                xyz = self.bc_centers(kind, i)
                vals = np.ascontiguousarray([bc["value"](kind, x) for x in xyz], dtype=np.float64)
                _check(self.L, self.L.hfh_bc_set_values(self.h, kind, i, _dp(vals)), "hfh_bc_set_values")
        self.ve = case["params"].get("flowtype", 0) == 3

    def close(self):
        if getattr(self, "h", None):
            self.L.hfh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bc_centers(self, kind, i):
        n = _check(self.L, self.L.hfh_bc_count(self.h, kind, i), "hfh_bc_count")
        out = np.zeros((n, self.dim))
        if n:
            self.L.hfh_bc_centers(self.h, kind, i, _dp(out))
        return out

    def set_bc(self, kind, i, values):
        """Boundary values of boundary tree i at its cell centres (kind 0
        pressure, 1+d velocity component d), as sb_set_value does."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.size != len(self.bc_centers(kind, i)):
            raise ValueError("boundary value count mismatch")
        _check(self.L, self.L.hfh_bc_set_values(self.h, kind, i, _dp(v)), "hfh_bc_set_values")

    def set_functions(self, **callables):
        """Register the driver callbacks (see make_functions) that every step
        re-evaluates on the host."""
        self._functions = make_functions(self.dim, **callables)
        _check(self.L, self.L.hfh_set_functions(self.h, C.byref(self._functions)), "hfh_set_functions")

    def set_valuetype(self, kind, i, timedependent):
        _check(self.L, self.L.hfh_bc_set_valuetype(self.h, kind, i, int(timedependent)), "hfh_bc_set_valuetype")

    def size(self, field, comp=0):
        return _check(self.L, self.L.hfh_field_size(self.h, field, comp), "hfh_field_size")

    def centers(self, field, comp=0):
        out = np.zeros((self.size(field, comp), self.dim))
        self.L.hfh_field_centers(self.h, field, comp, _dp(out))
        return out

    def get(self, field, comp=0):
        out = np.zeros(self.size(field, comp))
        _check(self.L, self.L.hfh_get_field(self.h, field, comp, _dp(out)), "hfh_get_field")
        return out

    def set(self, field, comp, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.size != self.size(field, comp):
            raise ValueError("field size mismatch")
        _check(self.L, self.L.hfh_set_field(self.h, field, comp, _dp(v)), "hfh_set_field")

    def range(self, field, comp=0):
        """(min, max) of a field, reduced on the device (for velocity: the
        Vmin/Vmax the reference's final-velocity step prints)."""
        mn, mx = C.c_double(0), C.c_double(0)
        _check(self.L, self.L.hfh_field_range(self.h, field, comp, C.byref(mn), C.byref(mx)), "hfh_field_range")
        return mn.value, mx.value

    def cell_bounds(self):
        """(lowpoints, highpoints) of the pressure cells, each (Nc, dim)."""
        n = self.size(1)
        out = np.zeros(2 * self.dim * n)
        _check(self.L, self.L.hfh_cell_bounds(self.h, _dp(out)), "hfh_cell_bounds")
        out = out.reshape(n, 2, self.dim)
        return out[:, 0, :].copy(), out[:, 1, :].copy()

    def step(self):
        if self.ve:
            return _check(self.L, self.L.hfh_solver_step_viscoelastic(self.h), "hfh_solver_step_viscoelastic")
        return _check(self.L, self.L.hfh_solver_step(self.h), "hfh_solver_step")

    def advance(self, n):
        rc = 0
        for _ in range(n):
            rc = max(rc, self.step())
        return rc

    def sync(self):
        _check(self.L, self.L.hfh_sync(self.h), "hfh_sync")

    def time(self):
        return self.L.hfh_get_time(self.h)

    def pinned_cell(self):
        return self.L.hfh_pinned_cell(self.h)

    def last_solve(self):
        it, outer = C.c_int(0), C.c_int(0)
        r = C.c_double(0)
        conv = self.L.hfh_last_solve(self.h, C.byref(it), C.byref(r), C.byref(outer))
        return conv, it.value, r.value, outer.value

    def solve_stats(self, which=0):
        """Last solve of a system (0 pressure, 1+d velocity component d):
        dict of converged, iters, outer, fallback, restarts, resid."""
        info = (C.c_int * 4)()
        r = C.c_double(0)
        conv = _check(self.L, self.L.hfh_solve_stats(self.h, which, info, C.byref(r)), "hfh_solve_stats")
        return dict(converged=conv, iters=info[0], outer=info[1], fallback=info[2], restarts=info[3],
                    resid=r.value)

    def stencil(self, domain, dim, center, x, alpha=1.0):
        ids = (C.c_int * 4096)()
        w = np.zeros(4096)
        rhs = C.c_double(0)
        c = np.ascontiguousarray(center, dtype=np.float64)
        xx = np.ascontiguousarray(x, dtype=np.float64)
        n = _check(self.L, self.L.hfh_stencil(self.h, domain, dim, _dp(c), _dp(xx), alpha, 4096, ids,
                                              _dp(w), C.byref(rhs)), "hfh_stencil")
        return np.array(ids[:n], dtype=np.int64), w[:n].copy(), rhs.value

    def timing(self):
        ms = np.zeros(len(STAGES))
        calls = (C.c_long * len(STAGES))()
        self.L.hfh_timing(self.h, _dp(ms), calls)
        return {s: (ms[i], calls[i]) for i, s in enumerate(STAGES)}

    def reset_timing(self):
        self.L.hfh_reset_timing(self.h)

    def set_timing(self, on):
        _check(self.L, self.L.hfh_set_timing(self.h, 1 if on else 0), "hfh_set_timing")

    def probe(self, which, reps):
        """(avg ms per launch, algorithmic bytes per launch) of one step kernel."""
        ms = C.c_double(0)
        by = C.c_long(0)
        rc = self.L.hfh_probe(self.h, which, reps, C.byref(ms), C.byref(by))
        if rc < 0:
            return None
        return ms.value, by.value

    def stage_bytes(self):
        return {s: self.L.hfh_stage_bytes(self.h, i) for i, s in enumerate(STAGES)}

    def tables_digest(self):
        """64-bit hash of the recorded host tables (same step recorded, same
        value)."""
        d = C.c_ulonglong(0)
        _check(self.L, self.L.hfh_tables_digest(self.h, C.byref(d)), "hfh_tables_digest")
        return d.value

    def operator(self, which=0):
        """(rowptr, cols, vals, symmetric) of an assembled system: 0 pressure,
        1+d the semi-implicit velocity system of component d."""
        n, nnz, sym = C.c_int(), C.c_int(), C.c_int()
        _check(self.L, self.L.hfh_operator_csr(self.h, which, C.byref(n), C.byref(nnz), C.byref(sym), None, None,
                                               None), "hfh_operator_csr")
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        v = np.zeros(nnz.value)
        _check(self.L, self.L.hfh_operator_csr(self.h, which, C.byref(n), C.byref(nnz), C.byref(sym), _ip(rp),
                                               _ip(ci), _dp(v)), "hfh_operator_csr")
        return rp, ci, v, bool(sym.value)


class Batch(Flow):
    """Independent simulations (case dicts of one dimension, same time step
    and schemes; Re, De and beta may differ) stepped together on the device:
    each step stage is one launch for all of them and the linear systems are
    solved block-diagonally with one stopping test per simulation.  Fields,
    boundary values and solve statistics are per member."""

    def __init__(self, case_list, **param_over):
        self.L = load()
        if not case_list:
            raise ValueError("empty batch")
        self.dim = case_list[0]["dim"]
        self.cases = list(case_list)
        # members set up side by side (the library releases the GIL; each
        # member's tables are its own)
        from concurrent.futures import ThreadPoolExecutor
        workers = max(1, min(16, os.cpu_count() or 1, len(self.cases)))
        with ThreadPoolExecutor(workers) as ex:
            members = list(ex.map(lambda c: Flow(c, host_only=True, **param_over), self.cases))
        arr = (C.c_void_p * len(members))(*[m.h for m in members])
        self.h = self.L.hfh_create_batch(self.dim, arr, len(members))
        if not self.h:
            raise HipError("hfh_create_batch: " + self.L.hfh_last_error().decode())
        for m in members:
            m.h = None  # freed by hfh_create_batch
        self.n = len(members)
        self.ve = self.cases[0]["params"].get("flowtype", 0) == 3
        self.params = members[0].params

    def size(self, member, field, comp=0):
        return _check(self.L, self.L.hfh_batch_field_size(self.h, member, field, comp), "hfh_batch_field_size")

    def get(self, member, field, comp=0):
        out = np.zeros(self.size(member, field, comp))
        _check(self.L, self.L.hfh_batch_get_field(self.h, member, field, comp, _dp(out)), "hfh_batch_get_field")
        return out

    def set(self, member, field, comp, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.size != self.size(member, field, comp):
            raise ValueError("field size mismatch")
        _check(self.L, self.L.hfh_batch_set_field(self.h, member, field, comp, _dp(v)), "hfh_batch_set_field")

    def solve_stats(self, member, which=0):
        info = (C.c_int * 4)()
        r = C.c_double(0)
        conv = _check(self.L, self.L.hfh_batch_solve_stats(self.h, member, which, info, C.byref(r)),
                      "hfh_batch_solve_stats")
        return dict(converged=conv, iters=info[0], outer=info[1], fallback=info[2], restarts=info[3],
                    resid=r.value)

    def member(self, m):
        """A view of member m with the single-run field interface."""
        return _Member(self, m)


class _Member:
    def __init__(self, batch, m):
        self.b, self.m = batch, m

    def size(self, field, comp=0):
        return self.b.size(self.m, field, comp)

    def get(self, field, comp=0):
        return self.b.get(self.m, field, comp)

    def set(self, field, comp, values):
        return self.b.set(self.m, field, comp, values)


class LinearSolver:
    """The HIP solver-plugin engine: the _solver_vtable members of
    higtree/src/solver.h:42-60 (slv_* semantics)."""

    def __init__(self, first_gid, size):
        self.L = load()
        self.h = self.L.hfh_slv_create(first_gid, size)
        if not self.h:
            raise HipError("hfh_slv_create: " + self.L.hfh_last_error().decode())
        self.n = size

    def close(self):
        if getattr(self, "h", None):
            self.L.hfh_slv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_Ai(self, i, cols, vals):
        cols = (C.c_int * len(cols))(*[int(c) for c in cols])
        v = np.ascontiguousarray(vals, dtype=np.float64)
        _check(self.L, self.L.hfh_slv_set_Ai(self.h, i, len(v), cols, _dp(v)), "hfh_slv_set_Ai")

    def set_bi(self, i, v):
        _check(self.L, self.L.hfh_slv_set_bi(self.h, i, v), "hfh_slv_set_bi")

    def add_bi(self, i, v):
        _check(self.L, self.L.hfh_slv_add_bi(self.h, i, v), "hfh_slv_add_bi")

    def set_xi(self, i, v):
        _check(self.L, self.L.hfh_slv_set_xi(self.h, i, v), "hfh_slv_set_xi")

    def impose_value(self, line, v):
        _check(self.L, self.L.hfh_slv_impose_value(self.h, line, v), "hfh_slv_impose_value")

    def set_tolerances(self, rtol, atol, maxit):
        _check(self.L, self.L.hfh_slv_set_tolerances(self.h, rtol, atol, maxit), "hfh_slv_set_tolerances")

    def assemble(self):
        for f in ("hfh_slv_assemble_rhs", "hfh_slv_assemble_output", "hfh_slv_assemble_matrix"):
            _check(self.L, getattr(self.L, f)(self.h), f)

    def solve(self):
        _check(self.L, self.L.hfh_slv_solve(self.h), "hfh_slv_solve")

    def get_x(self):
        out = np.zeros(self.n)
        _check(self.L, self.L.hfh_slv_get_x(self.h, _dp(out)), "hfh_slv_get_x")
        return out

    def stats(self):
        return (self.L.hfh_slv_converged(self.h), self.L.hfh_slv_iterations(self.h),
                self.L.hfh_slv_residual(self.h))

// C ABI of libhigflow_hip.so (declarations in include/higflow_hip.h).
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "../../include/higflow_hip.h"
#include "hfh_flow.hpp"

namespace {
thread_local std::string g_err;

template <class F>
int guard(F &&f) {
    try {
        g_err.clear();
        return f();
    } catch (const std::exception &e) {
        g_err = e.what();
    } catch (...) {
        g_err = "unknown error";
    }
    return -100;
}

int fail(const char *m) {
    g_err = m;
    return -1;
}

// Scratch of hfh_range_noalloc: the k_minmax ticket (first 256 bytes), then
// one (min, max) pair per workgroup for up to HFH_RANGE_NOALLOC_MAX elements.
// Each device holds its own copy for the life of the process; its address
// is looked up once per device with that device current and kept, since
// the array's name is not a device address in host code.
constexpr int RANGE_MAX_GROUPS = (int)(HFH_RANGE_NOALLOC_MAX / hfh::MM_CHUNK);
static_assert((long)RANGE_MAX_GROUPS * hfh::MM_CHUNK == HFH_RANGE_NOALLOC_MAX, "whole workgroups");
__device__ __align__(256) char scratch[256 + 2 * sizeof(double) * RANGE_MAX_GROUPS];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char *g_scratch[MAX_DEVICES] = {};

char *range_scratch() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        g_err = "hfh_range_noalloc: no current device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_err = std::string("hfh_range_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch[dev] = static_cast<char *>(p);
    }
    return g_scratch[dev];
}
}  // namespace

// --------------------------------------------------------------- engine
// Host-side row storage as the reference engines keep it (rows replaced by
// set_Ai, rhs/x vectors), device CSR built at assemble_matrix.
struct hfh_slv {
    int first = 0, n = 0;
    std::vector<std::vector<std::pair<int, double>>> rows;
    std::vector<double> b, x;
    int imposed = -1;
    bool dirty = true;
    double rtol = 1e-10, atol = 1e-10;
    int maxit = 10000;
    hfh::SolveStats st;
    hfh::DeviceSystem sys;
    hfh::DevBuf<double> db, dx;
    hipStream_t stream = nullptr;
    bool symmetric = false;
};

extern "C" {

const char *hfh_last_error(void) { return g_err.c_str(); }

int hfh_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int hfh_version(void) { return 100; }

hfh_slv *hfh_slv_create(int first_gid, int size) {
    if (size < 0 || first_gid < 0) {
        fail("hfh_slv_create: negative size or first id");
        return nullptr;
    }
    hfh_slv *s = nullptr;
    int rc = guard([&] {
        s = new hfh_slv();
        s->first = first_gid;
        s->n = size;
        s->rows.assign(size, {});
        s->b.assign(size, 0.0);
        s->x.assign(size, 0.0);
        HFH_CHECK(hipStreamCreate(&s->stream));
        return 0;
    });
    if (rc < 0) {
        delete s;
        return nullptr;
    }
    return s;
}

int hfh_slv_set_maxnonzeros(hfh_slv *s, int) { return s ? 0 : fail("null solver"); }

static int lrow(hfh_slv *s, int i) {
    if (!s) return -1;
    int r = i - s->first;
    if (r < 0 || r >= s->n) return -1;
    return r;
}

int hfh_slv_set_Ai(hfh_slv *s, int i, int n, const int *cols, const double *vals) {
    int r = lrow(s, i);
    if (r < 0 || n < 0) return fail("hfh_slv_set_Ai: row out of range");
    std::vector<std::pair<int, double>> row;
    for (int k = 0; k < n; k++) {
        int c = cols[k] - s->first;
        if (c < 0 || c >= s->n) return fail("hfh_slv_set_Ai: column out of range");
        row.emplace_back(c, vals[k]);
    }
    s->rows[r] = std::move(row);
    s->dirty = true;
    return 0;
}

int hfh_slv_set_bi(hfh_slv *s, int i, double v) {
    int r = lrow(s, i);
    if (r < 0) return fail("hfh_slv_set_bi: row out of range");
    if (s->imposed != r) s->b[r] = v;
    return 0;
}

int hfh_slv_add_bi(hfh_slv *s, int i, double v) {
    int r = lrow(s, i);
    if (r < 0) return fail("hfh_slv_add_bi: row out of range");
    if (s->imposed != r) s->b[r] += v;
    return 0;
}

int hfh_slv_set_xi(hfh_slv *s, int i, double v) {
    int r = lrow(s, i);
    if (r < 0) return fail("hfh_slv_set_xi: row out of range");
    s->x[r] = v;
    return 0;
}

int hfh_slv_set_x(hfh_slv *s, const double *x) {
    if (!s) return fail("null solver");
    for (int i = 0; i < s->n; i++) s->x[i] = x[i];
    return 0;
}

double hfh_slv_get_xi(hfh_slv *s, int i) {
    int r = lrow(s, i);
    if (r < 0) {
        fail("hfh_slv_get_xi: row out of range");
        return 0.0;
    }
    return s->x[r];
}

int hfh_slv_get_x(hfh_slv *s, double *x) {
    if (!s) return fail("null solver");
    for (int i = 0; i < s->n; i++) x[i] = s->x[i];
    return 0;
}

int hfh_slv_get_x_scatter(hfh_slv *s, int n, const int *idx, double *x) {
    if (!s) return fail("null solver");
    for (int k = 0; k < n; k++) {
        int r = lrow(s, idx[k]);
        if (r < 0) return fail("hfh_slv_get_x_scatter: index out of range");
        x[k] = s->x[r];
    }
    return 0;
}

int hfh_slv_impose_value(hfh_slv *s, int line, double value) {
    int r = lrow(s, line);
    if (r < 0) return fail("hfh_slv_impose_value: row out of range");
    s->imposed = r;
    s->b[r] = value;
    s->rows[r] = {{r, 1.0}};
    s->dirty = true;
    return 0;
}

int hfh_slv_assemble_matrix(hfh_slv *s) {
    if (!s) return fail("null solver");
    return guard([&] {
        if (!s->dirty) return 0;
        std::vector<int> rp(1, 0), ci;
        std::vector<double> v;
        std::vector<std::map<int, double>> m(s->n);
        for (int r = 0; r < s->n; r++)
            for (auto &e : s->rows[r]) m[r][e.first] += e.second;
        s->symmetric = true;
        for (int r = 0; r < s->n; r++) {
            for (auto &e : m[r]) {
                ci.push_back(e.first);
                v.push_back(e.second);
                if (s->symmetric) {
                    auto it = m[e.first].find(r);
                    if (it == m[e.first].end() || it->second != e.second) s->symmetric = false;
                }
            }
            rp.push_back((int)ci.size());
        }
        const char *ea = std::getenv("HFH_SLV_AMG");  // opt-in V-cycle preconditioning
        s->sys.set_matrix(rp, ci, v, s->stream, ea && std::strcmp(ea, "1") == 0);
        s->db.alloc(s->n);
        s->dx.alloc(s->n);
        s->dirty = false;
        return 0;
    });
}

int hfh_slv_assemble_rhs(hfh_slv *s) { return s ? 0 : fail("null solver"); }
int hfh_slv_assemble_output(hfh_slv *s) { return s ? 0 : fail("null solver"); }

int hfh_slv_solve(hfh_slv *s) {
    if (!s) return fail("null solver");
    return guard([&] {
        if (s->dirty) {
            int rc = hfh_slv_assemble_matrix(s);
            if (rc < 0) return rc;
        }
        if (s->n == 0) return 0;
        s->db.upload(s->b.data(), s->n, s->stream);
        s->dx.upload(s->x.data(), s->n, s->stream);
        s->st = s->sys.fused ? s->sys.fused64(false, s->db.p, s->dx.p, s->rtol, s->atol, s->maxit, s->stream)
                             : s->sys.bicgstab(s->db.p, s->dx.p, s->rtol, s->atol, s->maxit, s->stream);
        HFH_CHECK(hipMemcpyAsync(s->x.data(), s->dx.p, s->n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HFH_CHECK(hipStreamSynchronize(s->stream));
        return 0;
    });
}

void hfh_slv_destroy(hfh_slv *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int hfh_slv_set_tolerances(hfh_slv *s, double rtol, double atol, int maxit) {
    if (!s) return fail("null solver");
    s->rtol = rtol;
    s->atol = atol;
    s->maxit = maxit;
    return 0;
}

int hfh_slv_iterations(hfh_slv *s) { return s ? s->st.iters : fail("null solver"); }
double hfh_slv_residual(hfh_slv *s) { return s ? s->st.resid : -1.0; }
int hfh_slv_converged(hfh_slv *s) { return s ? s->st.converged : fail("null solver"); }

// ----------------------------------------------------------------- step
struct hfh_solver {
    std::unique_ptr<hfh::FlowBase> f;
    std::vector<std::unique_ptr<hfh::FlowBase>> members;  // batched runs
};

hfh_solver *hfh_create(int dim) {
    if (dim != 2 && dim != 3) {
        fail("hfh_create: dim must be 2 or 3");
        return nullptr;
    }
    hfh_solver *s = nullptr;
    guard([&] {
        const char *ho = std::getenv("HFH_HOST_ONLY");
        bool host_only = ho && ho[0] == '1';
        int n = 0;
        if (!host_only && (hipGetDeviceCount(&n) != hipSuccess || n == 0))
            throw std::runtime_error("no HIP device visible");
        s = new hfh_solver();
        if (dim == 2) s->f.reset(new hfh::Flow<2>(host_only));
        else s->f.reset(new hfh::Flow<3>(host_only));
        return 0;
    });
    return s;
}

hfh_solver *hfh_create_host(int dim) {
    if (dim != 2 && dim != 3) {
        fail("hfh_create_host: dim must be 2 or 3");
        return nullptr;
    }
    hfh_solver *s = new hfh_solver();
    if (dim == 2) s->f.reset(new hfh::Flow<2>(true));
    else s->f.reset(new hfh::Flow<3>(true));
    return s;
}

hfh_solver *hfh_create_batch(int dim, hfh_solver **members, int n) {
    if ((dim != 2 && dim != 3) || !members || n <= 0) {
        fail("hfh_create_batch: dim must be 2 or 3 and at least one member given");
        return nullptr;
    }
    for (int i = 0; i < n; i++)
        if (!members[i] || !members[i]->f || members[i]->f->dim() != dim) {
            fail("hfh_create_batch: null member or member of another dimension");
            return nullptr;
        }
    hfh_solver *s = nullptr;
    int rc = guard([&] {
        // HFH_HOST_ONLY=1: merge the host tables only (inspection, no device)
        const char *ho = std::getenv("HFH_HOST_ONLY");
        const bool host_only = ho && ho[0] == '1';
        int nd = 0;
        if (!host_only && (hipGetDeviceCount(&nd) != hipSuccess || nd == 0))
            throw std::runtime_error("no HIP device visible");
        std::unique_ptr<hfh_solver> b(new hfh_solver());
        if (dim == 2) b->f.reset(new hfh::Flow<2>(host_only));
        else b->f.reset(new hfh::Flow<3>(host_only));
        std::vector<hfh::FlowBase *> ms;
        for (int i = 0; i < n; i++) ms.push_back(members[i]->f.get());
        int r = b->f->setup_batch(ms);
        if (r < 0) throw std::runtime_error("hfh_create_batch: members not set up");
        for (int i = 0; i < n; i++) {
            b->members.push_back(std::move(members[i]->f));
            delete members[i];
            members[i] = nullptr;
        }
        s = b.release();
        return 0;
    });
    return rc < 0 ? nullptr : s;
}

void hfh_destroy(hfh_solver *s) { delete s; }

#define HFH_SOLVER(s)                                   \
    if (!s || !s->f) return fail("null hfh_solver");

int hfh_add_domain_amr(hfh_solver *s, const char *amr) {
    HFH_SOLVER(s);
    return guard([&] {
        int r = s->f->add_domain(amr ? amr : "");
        if (r < 0) g_err = "hfh_add_domain_amr: malformed .amr text or solver already set up";
        return r;
    });
}

int hfh_add_bc_amr(hfh_solver *s, int userid, const char *amr, int ptype, const int *utypes) {
    HFH_SOLVER(s);
    if (!amr || !utypes) return fail("hfh_add_bc_amr: null argument");
    hfh::BcSpec b;
    b.amr = amr;
    b.userid = userid;
    b.ptype = ptype;
    for (int d = 0; d < s->f->dim(); d++) b.utypes[d] = utypes[d];
    return guard([&] { return s->f->add_bc(b); });
}

int hfh_setup(hfh_solver *s, const hfh_params *p) {
    HFH_SOLVER(s);
    if (!p) return fail("hfh_setup: null params");
    if (p->dt <= 0 || p->Re <= 0) return fail("hfh_setup: dt and Re must be positive");
    hfh::Params q;
    q.Re = p->Re;
    q.dt = p->dt;
    q.t = p->t;
    q.projtype = p->projtype;
    q.flowtype = p->flowtype;
    q.tempdiscrtype = p->tempdiscrtype;
    q.convecdiscrtype = p->convecdiscrtype;
    q.secondconvecdiscrtype = p->secondconvecdiscrtype;
    q.ve_model = p->ve_model;
    q.ve_convec = p->ve_convec;
    q.kernel_transform = p->kernel_transform;
    q.De = p->De;
    q.beta = p->beta;
    q.kernel_tol = p->kernel_tol;
    q.giesekus_alpha = p->giesekus_alpha;
    q.rtol = p->rtol;
    q.atol = p->atol;
    q.maxit = p->maxit;
    q.pin_pressure = p->pin_pressure;
    q.order = p->order;
    q.precision = p->precision;
    q.inner_rtol = p->inner_rtol;
    return guard([&] {
        int r = s->f->setup(q);
        if (r < 0) g_err = "hfh_setup: no domain, already set up, or malformed boundary .amr";
        return r;
    });
}

int hfh_bc_count(hfh_solver *s, int kind, int i) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->bc_count(kind, i); });
}
int hfh_bc_centers(hfh_solver *s, int kind, int i, double *xyz) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->bc_centers(kind, i, xyz); });
}
int hfh_bc_set_values(hfh_solver *s, int kind, int i, const double *v) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->bc_set_values(kind, i, v); });
}
int hfh_set_functions(hfh_solver *s, const hfh_functions *f) {
    HFH_SOLVER(s);
    hfh::Functions fn;
    if (f) {
        fn.get_boundary_pressure = f->get_boundary_pressure;
        fn.get_boundary_velocity = f->get_boundary_velocity;
        fn.get_source_term = f->get_source_term;
        fn.get_facet_source_term = f->get_facet_source_term;
    }
    return guard([&] {
        int r = s->f->set_functions(fn);
        if (r < 0) g_err = "hfh_set_functions: not set up, or a batch";
        return r;
    });
}
int hfh_bc_set_valuetype(hfh_solver *s, int kind, int i, int timedependent) {
    HFH_SOLVER(s);
    return guard([&] {
        int r = s->f->bc_set_valuetype(kind, i, timedependent);
        if (r < 0) g_err = "hfh_bc_set_valuetype: no such boundary, not set up, or a batch";
        return r;
    });
}
int hfh_field_size(hfh_solver *s, int field, int comp) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->field_size(field, comp); });
}
int hfh_field_centers(hfh_solver *s, int field, int comp, double *xyz) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->field_centers(field, comp, xyz); });
}
int hfh_get_field(hfh_solver *s, int field, int comp, double *out) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->get_field(field, comp, out); });
}
int hfh_cell_bounds(hfh_solver *s, double *lohi) {
    HFH_SOLVER(s);
    if (!lohi) {
        g_err = "hfh_cell_bounds: null output";
        return -1;
    }
    return guard([&] { return s->f->cell_bounds(lohi); });
}
int hfh_field_range(hfh_solver *s, int field, int comp, double *vmin, double *vmax) {
    HFH_SOLVER(s);
    if (!vmin || !vmax) {
        g_err = "hfh_field_range: null output";
        return -1;
    }
    return guard([&] { return s->f->field_range(field, comp, vmin, vmax); });
}
int hfh_set_field(hfh_solver *s, int field, int comp, const double *in) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->set_field(field, comp, in); });
}
int hfh_solver_step(hfh_solver *s) {
    HFH_SOLVER(s);
    return guard([&] {
        int r = s->f->step(false);
        if (r == -3) g_err = "hfh_solver_step: only explicit Euler is built";
        if (r == -4) g_err = "hfh_solver_step: viscoelastic case, use hfh_solver_step_viscoelastic";
        return r;
    });
}
int hfh_solver_step_viscoelastic(hfh_solver *s) {
    HFH_SOLVER(s);
    return guard([&] {
        int r = s->f->step(true);
        if (r == -3) g_err = "hfh_solver_step_viscoelastic: only explicit Euler is built";
        if (r == -4) g_err = "hfh_solver_step_viscoelastic: case is not viscoelastic";
        return r;
    });
}
int hfh_set_time(hfh_solver *s, double t) {
    HFH_SOLVER(s);
    s->f->set_time(t);
    return 0;
}
double hfh_get_time(hfh_solver *s) { return (s && s->f) ? s->f->time() : 0.0; }
int hfh_pinned_cell(hfh_solver *s) {
    HFH_SOLVER(s);
    return s->f->pinned();
}
int hfh_solve_stats(hfh_solver *s, int which, int *info, double *resid) {
    HFH_SOLVER(s);
    if (!info || !resid) return fail("null output pointers");
    hfh::SolveStats st;
    if (s->f->solve_stats(which, &st) < 0) return fail("no such system");
    info[0] = st.iters;
    info[1] = st.outer;
    info[2] = st.fallback;
    info[3] = st.restarts;
    *resid = st.resid;
    return st.converged ? 1 : 0;
}
int hfh_batch_size(hfh_solver *s) {
    HFH_SOLVER(s);
    return s->f->batch_size();
}
int hfh_batch_field_size(hfh_solver *s, int member, int field, int comp) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->member_field(member, field, comp, nullptr, false); });
}
int hfh_batch_get_field(hfh_solver *s, int member, int field, int comp, double *out) {
    HFH_SOLVER(s);
    if (!out) return fail("null output");
    return guard([&] { return s->f->member_field(member, field, comp, out, false); });
}
int hfh_batch_set_field(hfh_solver *s, int member, int field, int comp, const double *in) {
    HFH_SOLVER(s);
    if (!in) return fail("null input");
    return guard([&] { return s->f->member_field(member, field, comp, const_cast<double *>(in), true); });
}
int hfh_batch_bc_set_values(hfh_solver *s, int member, int kind, int i, const double *vals) {
    HFH_SOLVER(s);
    if (!vals) return fail("null values");
    return guard([&] { return s->f->member_bc_set_values(member, kind, i, vals); });
}
int hfh_batch_solve_stats(hfh_solver *s, int member, int which, int *info, double *resid) {
    HFH_SOLVER(s);
    if (!info || !resid) return fail("null output pointers");
    hfh::SolveStats st;
    if (s->f->member_stats(member, which, &st) < 0) return fail("no such member or system");
    info[0] = st.iters;
    info[1] = st.outer;
    info[2] = st.fallback;
    info[3] = st.restarts;
    *resid = st.resid;
    return st.converged ? 1 : 0;
}
int hfh_last_solve(hfh_solver *s, int *iters, double *resid, int *outer) {
    HFH_SOLVER(s);
    hfh::SolveStats st = s->f->last_solve();
    if (iters) *iters = st.iters;
    if (resid) *resid = st.resid;
    if (outer) *outer = st.outer;
    return st.converged;
}
int hfh_stencil(hfh_solver *s, int domain, int dim, const double *center, const double *x, double alpha, int maxn,
                int *ids, double *w, double *rhs) {
    HFH_SOLVER(s);
    return guard([&] { return s->f->stencil(domain, dim, center, x, alpha, maxn, ids, w, rhs); });
}
int hfh_timing(hfh_solver *s, double *ms, long *calls) {
    HFH_SOLVER(s);
    s->f->timing(ms, calls);
    return 0;
}
int hfh_reset_timing(hfh_solver *s) {
    HFH_SOLVER(s);
    s->f->reset_timing();
    return 0;
}
int hfh_set_timing(hfh_solver *s, int on) {
    HFH_SOLVER(s);
    s->f->set_timing(on != 0);
    return 0;
}
int hfh_probe(hfh_solver *s, int which, int reps, double *avg_ms, long *bytes) {
    HFH_SOLVER(s);
    return guard([&] {
        int r = s->f->probe(which, reps, avg_ms, bytes);
        if (r < 0) g_err = "hfh_probe: unknown kernel id for this case, or no device";
        return r;
    });
}
int hfh_operator_csr(hfh_solver *s, int which, int *n, int *nnz, int *sym, int *rp, int *ci, double *v) {
    HFH_SOLVER(s);
    if (!n || !nnz || !sym) return fail("null size pointers");
    return guard([&] {
        if (s->f->operator_csr(which, n, nnz, sym, rp, ci, v) < 0) return fail("no such operator");
        return 0;
    });
}
long hfh_stage_bytes(hfh_solver *s, int stage) { return (s && s->f) ? s->f->bytes(stage) : -1; }
int hfh_tables_digest(hfh_solver *s, unsigned long long *out) {
    HFH_SOLVER(s);
    if (!out) return fail("null digest pointer");
    return guard([&] {
        if (s->f->tables_digest(out) < 0) return fail("no recorded tables (call hfh_setup first)");
        return 0;
    });
}

int hfh_range_noalloc(const double *x, long n, double *minmax, void *stream) {
    if (n < 0 || n > HFH_RANGE_NOALLOC_MAX)
        return fail("hfh_range_noalloc: n is negative or above HFH_RANGE_NOALLOC_MAX (33,554,432)");
    if (!minmax || (n > 0 && !x)) return fail("hfh_range_noalloc: null device pointer");
    return guard([&] {
        char *sp = range_scratch();
        if (!sp) return -1;
        unsigned *ticket = reinterpret_cast<unsigned *>(sp);
        double *part = reinterpret_cast<double *>(sp + 256);
        hipStream_t st = static_cast<hipStream_t>(stream);
        hfh::k_zero_u32<<<1, 64, 0, st>>>(ticket);
        hfh::k_minmax<<<hfh::mm_groups(n), 1024, 0, st>>>(n, x, part, ticket, minmax);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(hipGetErrorString(e));
        return 0;
    });
}
int hfh_sync(hfh_solver *s) {
    HFH_SOLVER(s);
    return guard([&] {
        s->f->sync();
        return 0;
    });
}

}  // extern "C"

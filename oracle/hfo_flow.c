/*
 * CPU oracle (test infrastructure): the HiG-Flow projection step.
 * Restates higflow/src/hig-flow-step.c, hig-flow-discret.c, hig-flow-eval.c,
 * hig-flow-terms.c, hig-flow-step-viscoelastic.c,
 * hig-flow-step-generalized-newtonian.c and the PETSc KSP semantics of
 * higtree/src/solver-petsc.c (BiCGStab, right ILU(0), unpreconditioned
 * residual, warm start).  See hfo_internal.h for the oracle's role.
 */
#include "hfo_internal.h"
#include "hfo.h"

enum { NON_INCREMENTAL = 0, INCREMENTAL = 1 };
enum { NEWTONIAN = 0, VISCOELASTIC = 3 };
enum { EXPLICIT_EULER = 0, SEMI_IMPLICIT_EULER = 3 };
enum { CENTRAL = 0, FIRST_ORDER = 1, SECOND_ORDER = 2 };
enum { MCU = 0, CUBISTA = 1, QUICK = 2 };
enum { OLDROYD_B = 0, GIESEKUS = 1 };
enum { CELL_CENTRAL = 0, CELL_CUBISTA = 1 };

typedef struct {
    char *text;
    int userid, ptype;
    int utypes[DIM];
} bcspec;

/* sparse system of the solver plugin (higtree/src/solver.h:62-88) */
typedef struct {
    int n;
    int *rowcnt;
    int **rcols;
    real **rvals;
    real *b, *x;
    int imposed;
    int iters;
    real resid;
    int converged;
} lsys;

/* higflow_compcell (higflow/src/hig-flow-kernel.h:34-105), ORDER2 subset */
typedef struct {
    int convec_type;
    real ufacet, F, dpdx;
    real v[DIM], vc[DIM], dudxc[DIM], dudxr[DIM], dudxl[DIM];
    real dudxrr[DIM], dudxll[DIM], du2dx2[DIM], dSdx[DIM];
} compcell;

struct hfo_solver {
    idgen ids;
    hcell **trees;
    int ntrees;
    bcspec *specs;
    int nspecs;
    hfo_params par;
    int ready;
    bool desingpressure;
    hdomain *sdp, *sdF, *sdED;
    hfacetdomain *sfdu[DIM], *sfdF[DIM];
    hcell **bctrees_p, **bctrees_u;
    hboundary **bc_p;          /* per spec */
    hboundary **bc_u[DIM];     /* per spec */
    real *u[DIM], *ustar[DIM], *p, *dp, *F, *FU[DIM];
    real *K[DIM][DIM], *S[DIM][DIM], *D[DIM][DIM];
    hstencil stn;
    compcell cc;
    lsys ps;
    lsys us[DIM]; /* semi-implicit velocity systems (slvu[dim]) */
    int pin;
    hfo_functions fn;          /* higflow_functions callbacks, or zeros */
    char *timedep;             /* [(DIM + 1) * nspecs]: kind-major valuetype */
};

/* ------------------------------------------------------------ solver */
static void lsys_init(lsys *L, int n)
{
    L->n = n;
    L->rowcnt = calloc((size_t)n, sizeof(int));
    L->rcols = calloc((size_t)n, sizeof(int *));
    L->rvals = calloc((size_t)n, sizeof(real *));
    L->b = calloc((size_t)n, sizeof(real));
    L->x = calloc((size_t)n, sizeof(real));
    L->imposed = -1;
}

static void lsys_free(lsys *L)
{
    for (int i = 0; i < L->n; i++) {
        free(L->rcols[i]);
        free(L->rvals[i]);
    }
    free(L->rowcnt);
    free(L->rcols);
    free(L->rvals);
    free(L->b);
    free(L->x);
}

static void lsys_set_row(lsys *L, int i, int n, const int *cols, const real *vals)
{
    free(L->rcols[i]);
    free(L->rvals[i]);
    L->rcols[i] = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    L->rvals[i] = malloc(sizeof(real) * (size_t)(n > 0 ? n : 1));
    /* insertion sort by column (PETSc stores rows sorted) */
    int m = 0;
    for (int k = 0; k < n; k++) {
        int c = cols[k];
        real v = vals[k];
        int q = m;
        while (q > 0 && L->rcols[i][q - 1] > c) {
            L->rcols[i][q] = L->rcols[i][q - 1];
            L->rvals[i][q] = L->rvals[i][q - 1];
            q--;
        }
        L->rcols[i][q] = c;
        L->rvals[i][q] = v;
        m++;
    }
    L->rowcnt[i] = n;
}

/* higtree/src/solver.c:142-149 */
static void lsys_impose(lsys *L, int line, real value)
{
    L->imposed = line;
    L->b[line] = value;
    const real one = 1.0;
    lsys_set_row(L, line, 1, &line, &one);
}

/* higtree/src/solver.c:99-105 */
static void lsys_set_bi(lsys *L, int i, real v)
{
    if (L->imposed != i) L->b[i] = v;
}

typedef struct {
    int n, nnz;
    int *rp, *ci;
    real *v;
    int *diag;
    real *lu;
} csr;

static void build_csr(lsys *L, csr *A)
{
    A->n = L->n;
    A->rp = malloc(sizeof(int) * (size_t)(L->n + 1));
    A->rp[0] = 0;
    for (int i = 0; i < L->n; i++) A->rp[i + 1] = A->rp[i] + L->rowcnt[i];
    A->nnz = A->rp[L->n];
    A->ci = malloc(sizeof(int) * (size_t)(A->nnz + 1));
    A->v = malloc(sizeof(real) * (size_t)(A->nnz + 1));
    A->diag = malloc(sizeof(int) * (size_t)(L->n + 1));
    for (int i = 0; i < L->n; i++) {
        A->diag[i] = -1;
        for (int k = 0; k < L->rowcnt[i]; k++) {
            A->ci[A->rp[i] + k] = L->rcols[i][k];
            A->v[A->rp[i] + k] = L->rvals[i][k];
            if (L->rcols[i][k] == i) A->diag[i] = A->rp[i] + k;
        }
    }
    /* ILU(0): PETSc -pc_type bjacobi on one rank = ILU(0) of the whole matrix */
    A->lu = malloc(sizeof(real) * (size_t)(A->nnz + 1));
    memcpy(A->lu, A->v, sizeof(real) * (size_t)A->nnz);
    for (int i = 0; i < A->n; i++) {
        for (int kk = A->rp[i]; kk < A->rp[i + 1] && A->ci[kk] < i; kk++) {
            int k = A->ci[kk];
            if (A->diag[k] < 0) continue;
            real piv = A->lu[kk] / A->lu[A->diag[k]];
            A->lu[kk] = piv;
            int jj = kk + 1, kj = A->diag[k] + 1;
            while (jj < A->rp[i + 1] && kj < A->rp[k + 1]) {
                if (A->ci[jj] == A->ci[kj]) {
                    A->lu[jj] -= piv * A->lu[kj];
                    jj++;
                    kj++;
                } else if (A->ci[jj] < A->ci[kj]) {
                    jj++;
                } else {
                    kj++;
                }
            }
        }
    }
}

static void free_csr(csr *A)
{
    free(A->rp);
    free(A->ci);
    free(A->v);
    free(A->diag);
    free(A->lu);
}

static void spmv(const csr *A, const real *x, real *y)
{
    for (int i = 0; i < A->n; i++) {
        real s = 0.0;
        for (int k = A->rp[i]; k < A->rp[i + 1]; k++) s += A->v[k] * x[A->ci[k]];
        y[i] = s;
    }
}

static void ilu_apply(const csr *A, const real *r, real *z)
{
    for (int i = 0; i < A->n; i++) {
        real s = r[i];
        for (int k = A->rp[i]; k < A->rp[i + 1] && A->ci[k] < i; k++) s -= A->lu[k] * z[A->ci[k]];
        z[i] = s;
    }
    for (int i = A->n - 1; i >= 0; i--) {
        real s = z[i];
        for (int k = A->diag[i] + 1; k < A->rp[i + 1]; k++) s -= A->lu[k] * z[A->ci[k]];
        z[i] = s / A->lu[A->diag[i]];
    }
}

static real dot(int n, const real *a, const real *b)
{
    real s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* BiCGStab (van der Vorst 1992), right preconditioned, convergence on the
 * true residual ||b - Ax|| <= max(rtol ||b||, atol) as PETSc's default test
 * (higtree/src/solver-petsc.c:387-394 with -ksp_type bcgs -pc_type bjacobi). */
static void lsys_solve(lsys *L, real rtol, real atol, int maxit)
{
    csr A;
    build_csr(L, &A);
    int n = L->n;
    real *r = calloc((size_t)n, sizeof(real)), *rh = calloc((size_t)n, sizeof(real));
    real *pv = calloc((size_t)n, sizeof(real)), *v = calloc((size_t)n, sizeof(real));
    real *ph = calloc((size_t)n, sizeof(real)), *sv = calloc((size_t)n, sizeof(real));
    real *sh = calloc((size_t)n, sizeof(real)), *t = calloc((size_t)n, sizeof(real));
    real *x = L->x;
    spmv(&A, x, r);
    for (int i = 0; i < n; i++) r[i] = L->b[i] - r[i];
    memcpy(rh, r, sizeof(real) * (size_t)n);
    real bnorm = sqrt(dot(n, L->b, L->b));
    real tol = rtol * bnorm;
    if (tol < atol) tol = atol;
    real rnorm = sqrt(dot(n, r, r));
    real rho = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0;
    L->converged = rnorm <= tol;
    while (!L->converged && it < maxit) {
        real rho1 = dot(n, rh, r);
        if (rho1 == 0.0) break;
        if (it == 0) {
            memcpy(pv, r, sizeof(real) * (size_t)n);
        } else {
            real beta = (rho1 / rho) * (alpha / omega);
            for (int i = 0; i < n; i++) pv[i] = r[i] + beta * (pv[i] - omega * v[i]);
        }
        ilu_apply(&A, pv, ph);
        spmv(&A, ph, v);
        real rv = dot(n, rh, v);
        if (rv == 0.0) break;
        alpha = rho1 / rv;
        for (int i = 0; i < n; i++) sv[i] = r[i] - alpha * v[i];
        it++;
        real snorm = sqrt(dot(n, sv, sv));
        if (snorm <= tol) {
            for (int i = 0; i < n; i++) x[i] += alpha * ph[i];
            rnorm = snorm;
            L->converged = 1;
            break;
        }
        ilu_apply(&A, sv, sh);
        spmv(&A, sh, t);
        real tt = dot(n, t, t);
        omega = tt > 0.0 ? dot(n, t, sv) / tt : 0.0;
        for (int i = 0; i < n; i++) {
            x[i] += alpha * ph[i] + omega * sh[i];
            r[i] = sv[i] - omega * t[i];
        }
        rnorm = sqrt(dot(n, r, r));
        rho = rho1;
        if (rnorm <= tol) L->converged = 1;
        if (omega == 0.0) break;
    }
    L->iters = it;
    L->resid = bnorm > 0 ? rnorm / bnorm : rnorm;
    free(r);
    free(rh);
    free(pv);
    free(v);
    free(ph);
    free(sv);
    free(sh);
    free(t);
    free_csr(&A);
}

/* ------------------------------------------------ point evaluators */
/* higflow/src/hig-flow-eval.c:43-58 */
static real eval_facet(hfacetdomain *f, const real *center, const real *p, const real *vals,
                       hstencil *stn)
{
    stn_reset(stn);
    hfo_sfd_get_stencil(f, center, p, 1.0, stn);
    return hfo_interpolate(vals, stn);
}

/* higflow/src/hig-flow-eval.c:403-412 */
static real eval_cell(hdomain *d, const real *center, const real *p, const real *vals,
                      hstencil *stn)
{
    stn_reset(stn);
    hfo_sd_get_stencil(d, center, p, 1.0, stn);
    return hfo_interpolate(vals, stn);
}

/* higflow/src/hig-flow-eval.c:199-210 */
static real facet_u_left(hfacetdomain *f, const real *c, const real *delta, int dim,
                         real alpha, const real *vals, hstencil *stn, int *infacet)
{
    Point p;
    hfacet fx;
    memcpy(p, c, sizeof p);
    p[dim] = c[dim] - alpha * delta[dim];
    *infacet = hfo_sfd_facet_with_point(f, p, &fx);
    return eval_facet(f, c, p, vals, stn);
}

/* higflow/src/hig-flow-eval.c:274-284 */
static real facet_u_right(hfacetdomain *f, const real *c, const real *delta, int dim,
                          real alpha, const real *vals, hstencil *stn, int *infacet)
{
    Point p;
    hfacet fx;
    memcpy(p, c, sizeof p);
    p[dim] = c[dim] + alpha * delta[dim];
    *infacet = hfo_sfd_facet_with_point(f, p, &fx);
    return eval_facet(f, c, p, vals, stn);
}

/* higflow/src/hig-flow-eval.c:60-92 */
static real facet_u_2(hfacetdomain *f, const real *c, const real *delta, int dim, int dim2,
                      real alpha, const real *vals, hstencil *stn)
{
    Point p;
    memcpy(p, c, sizeof p);
    p[dim] = c[dim] + 0.5 * alpha * delta[dim];
    p[dim2] = c[dim2] + 0.5 * alpha * delta[dim2];
    real vrr = eval_facet(f, c, p, vals, stn);
    p[dim] = c[dim] + 0.5 * alpha * delta[dim];
    p[dim2] = c[dim2] - 0.5 * alpha * delta[dim2];
    real vlr = eval_facet(f, c, p, vals, stn);
    p[dim] = c[dim] - 0.5 * alpha * delta[dim];
    p[dim2] = c[dim2] + 0.5 * alpha * delta[dim2];
    real vrl = eval_facet(f, c, p, vals, stn);
    p[dim] = c[dim] - 0.5 * alpha * delta[dim];
    p[dim2] = c[dim2] - 0.5 * alpha * delta[dim2];
    real vll = eval_facet(f, c, p, vals, stn);
    return 0.25 * (vll + vlr + vrl + vrr);
}

/* higflow/src/hig-flow-eval.c:95-121 (sgn=+1) and :160-186 (sgn=-1) */
static real facet_u_4(hfacetdomain *f, const real *c, const real *delta, int dim, int dim2,
                      real alpha, real sgn, const real *vals, hstencil *stn)
{
    Point p;
    memcpy(p, c, sizeof p);
    p[dim] = c[dim] + 0.5 * alpha * delta[dim];
    p[dim2] = (sgn > 0) ? c[dim2] + alpha * delta[dim2] : c[dim2] - alpha * delta[dim2];
    real vrr = eval_facet(f, c, p, vals, stn);
    p[dim] = c[dim] + 0.5 * alpha * delta[dim];
    p[dim2] = c[dim2];
    real vr = eval_facet(f, c, p, vals, stn);
    p[dim] = c[dim] - 0.5 * alpha * delta[dim];
    p[dim2] = c[dim2];
    real vl = eval_facet(f, c, p, vals, stn);
    p[dim] = c[dim] - 0.5 * alpha * delta[dim];
    p[dim2] = (sgn > 0) ? c[dim2] + alpha * delta[dim2] : c[dim2] - alpha * delta[dim2];
    real vll = eval_facet(f, c, p, vals, stn);
    return 0.25 * (vll + vl + vr + vrr);
}

/* higflow/src/hig-flow-eval.c:415-437 */
static real center_p_left(hdomain *d, const real *c, const real *delta, int dim, real alpha,
                          const real *vals, hstencil *stn)
{
    Point p;
    memcpy(p, c, sizeof p);
    p[dim] = c[dim] - alpha * delta[dim];
    return eval_cell(d, c, p, vals, stn);
}

static real center_p_right(hdomain *d, const real *c, const real *delta, int dim, real alpha,
                           const real *vals, hstencil *stn)
{
    Point p;
    memcpy(p, c, sizeof p);
    p[dim] = c[dim] + alpha * delta[dim];
    return eval_cell(d, c, p, vals, stn);
}

/* higflow/src/hig-flow-eval.c:440-474 (sgn=-1 left_2, +1 right_2) */
static real center_p_2(hdomain *d, const real *c, const real *delta, int dim, int dim2,
                       real alpha, real sgn, const real *vals, hstencil *stn)
{
    Point p;
    memcpy(p, c, sizeof p);
    p[dim2] = (sgn > 0) ? c[dim2] + alpha * delta[dim2] : c[dim2] - alpha * delta[dim2];
    p[dim] = c[dim] - 0.5 * alpha * delta[dim];
    real vl = eval_cell(d, c, p, vals, stn);
    p[dim2] = (sgn > 0) ? c[dim2] + alpha * delta[dim2] : c[dim2] - alpha * delta[dim2];
    p[dim] = c[dim] + 0.5 * alpha * delta[dim];
    real vr = eval_cell(d, c, p, vals, stn);
    return 0.5 * (vl + vr);
}

/* higflow/src/hig-flow-eval.c:476-500 */
static real center_p_22(hdomain *d, const real *c, const real *delta, int dim, real alpha,
                        real sgn, const real *vals, hstencil *stn, int *incell)
{
    Point p;
    memcpy(p, c, sizeof p);
    p[dim] = (sgn > 0) ? c[dim] + alpha * delta[dim] : c[dim] - alpha * delta[dim];
    *incell = hfo_sd_cell_with_point(d, p) != NULL;
    return eval_cell(d, c, p, vals, stn);
}

/* higflow/src/hig-flow-discret.c:12-79 */
static real dudxc(const real *delta, int dim, real alpha, real uc, real ul, real ur)
{
    (void)uc;
    return (ur - ul) / (2.0 * alpha * delta[dim]);
}
static real dudxr(const real *delta, int dim, real alpha, real uc, real ul, real ur)
{
    (void)ul;
    return (ur - uc) / (alpha * delta[dim]);
}
static real dudxl(const real *delta, int dim, real alpha, real uc, real ul, real ur)
{
    (void)ur;
    return (uc - ul) / (alpha * delta[dim]);
}
static real du2dx2(const real *delta, int dim, real alpha, real u, real ul, real ur)
{
    return (ul - 2.0 * u + ur) / (delta[dim] * delta[dim] * alpha * alpha);
}
static real dpdx_at(const real *delta, int dim, real alpha, real vl, real vh)
{
    return (vh - vl) / (2.0 * alpha * delta[dim]);
}
static real dpdxr_at(const real *delta, int dim, real alpha, real vc, real vr)
{
    return (vr - vc) / (alpha * delta[dim]);
}
static real dpdxl_at(const real *delta, int dim, real alpha, real vl, real vc)
{
    return (vc - vl) / (alpha * delta[dim]);
}
static real mid(real l, real r) { return (r + l) / 2.0; }

/* ----------------------------------------------- computational cell */
/* higflow/src/hig-flow-discret.c:86-259 (ORDER2) */
static void computational_cell(hfo_solver *s, int flid, const real *fc, const real *fd,
                               int dim, real *const *dpu)
{
    compcell *cc = &s->cc;
    hstencil *stn = &s->stn;
    hfacetdomain **sfdu = s->sfdu;
    real ul[DIM], ur[DIM], ull[DIM], urr[DIM], pl, pr, Fl, Fr;
    const real a = 1.7500, b = 0.3750, c = 0.7500, d = 0.1250, e = 0.2500;
    real fi, conv1 = 0.0, conv2 = 0.0;
    int infacet, infacet_r, infacet_l, infacet_rr, infacet_ll;
    cc->convec_type = s->par.convecdiscrtype;
    cc->ufacet = dpu[dim][flid];
    cc->F = eval_facet(s->sfdF[dim], fc, fc, s->FU[dim], stn);
    pl = center_p_left(s->sdp, fc, fd, dim, 0.5, s->p, stn);
    pr = center_p_right(s->sdp, fc, fd, dim, 0.5, s->p, stn);
    cc->dpdx = dpdx_at(fd, dim, 0.5, pl, pr);
    Fl = center_p_left(s->sdF, fc, fd, dim, 0.5, s->F, stn);
    Fr = center_p_right(s->sdF, fc, fd, dim, 0.5, s->F, stn);
    cc->F += mid(Fl, Fr);
    for (int dim2 = 0; dim2 < DIM; dim2++) {
        if (dim2 == dim) cc->v[dim2] = cc->ufacet;
        else cc->v[dim2] = facet_u_2(sfdu[dim2], fc, fd, dim, dim2, 1.0, dpu[dim2], stn);
        ul[dim2] = facet_u_left(sfdu[dim], fc, fd, dim2, 1.0, dpu[dim], stn, &infacet_l);
        if (infacet_l == 0 && cc->convec_type == SECOND_ORDER) cc->convec_type = FIRST_ORDER;
        ur[dim2] = facet_u_right(sfdu[dim], fc, fd, dim2, 1.0, dpu[dim], stn, &infacet_r);
        if (infacet_r == 0 && cc->convec_type == SECOND_ORDER) cc->convec_type = FIRST_ORDER;
        cc->dudxc[dim2] = dudxc(fd, dim2, 1.0, cc->ufacet, ul[dim2], ur[dim2]);
        cc->dudxr[dim2] = dudxr(fd, dim2, 1.0, cc->ufacet, ul[dim2], ur[dim2]);
        cc->dudxl[dim2] = dudxl(fd, dim2, 1.0, cc->ufacet, ul[dim2], ur[dim2]);
        urr[dim2] = facet_u_right(sfdu[dim], fc, fd, dim2, 2.0, dpu[dim], stn, &infacet_rr);
        ull[dim2] = facet_u_left(sfdu[dim], fc, fd, dim2, 2.0, dpu[dim], stn, &infacet_ll);
        if (cc->convec_type == SECOND_ORDER) {
            if (s->par.secondconvecdiscrtype == QUICK) {
                cc->dudxrr[dim2] = dudxr(fd, dim2, 1.0, ur[dim2], cc->ufacet, urr[dim2]);
                cc->dudxll[dim2] = dudxl(fd, dim2, 1.0, ul[dim2], ull[dim2], cc->ufacet);
            } else if (s->par.secondconvecdiscrtype == CUBISTA) {
                ul[dim] = facet_u_left(sfdu[dim], fc, fd, dim2, 1.0, dpu[dim], stn, &infacet_l);
                ull[dim] = facet_u_left(sfdu[dim], fc, fd, dim2, 2.0, dpu[dim], stn, &infacet_ll);
                if (infacet_l == 0) cc->convec_type = FIRST_ORDER;
                ur[dim] = facet_u_right(sfdu[dim], fc, fd, dim2, 1.0, dpu[dim], stn, &infacet_r);
                urr[dim] = facet_u_right(sfdu[dim], fc, fd, dim2, 2.0, dpu[dim], stn, &infacet_rr);
                if (infacet_r == 0) cc->convec_type = FIRST_ORDER;
                real uf = cc->ufacet;
                cc->vc[dim2] = facet_u_right(sfdu[dim2], fc, fd, dim2, 0.5, dpu[dim2], stn, &infacet);
                real vc = cc->vc[dim2];
                if (vc > 0.0) {
                    if (FLT_EQ(ur[dim], ul[dim])) {
                        conv1 = vc * uf;
                    } else {
                        fi = (uf - ul[dim]) / (ur[dim] - ul[dim]);
                        if ((fi <= 0.0) || (fi >= 1.0)) {
                            conv1 = vc * uf;
                        } else {
                            if (fi < b) conv1 = vc * (a * uf - c * ul[dim]);
                            if ((fi >= b) && (fi <= c)) conv1 = vc * (c * uf + b * ur[dim] - d * ul[dim]);
                            if (fi > c) conv1 = vc * (e * uf + c * ur[dim]);
                        }
                    }
                } else {
                    if (FLT_EQ(uf, urr[dim])) {
                        conv1 = vc * ur[dim];
                    } else {
                        fi = (ur[dim] - urr[dim]) / (uf - urr[dim]);
                        if ((fi <= 0.0) || (fi >= 1.0)) {
                            conv1 = vc * ur[dim];
                        } else {
                            if (fi < b) {
                                if (infacet_rr == 1) conv1 = vc * (a * ur[dim] - c * urr[dim]);
                                else conv1 = vc * ur[dim];
                            }
                            if ((fi >= b) && (fi <= c)) {
                                if (infacet_rr == 1) conv1 = vc * (c * ur[dim] + b * uf - d * urr[dim]);
                                else conv1 = vc * ur[dim];
                            }
                            if (fi > c) conv1 = vc * (c * uf + e * ur[dim]);
                        }
                    }
                }
                cc->vc[dim2] = facet_u_left(sfdu[dim2], fc, fd, dim2, 0.5, dpu[dim2], stn, &infacet);
                vc = cc->vc[dim2];
                if (vc > 0.0) {
                    if (FLT_EQ(uf, ull[dim])) {
                        conv2 = vc * ul[dim];
                    } else {
                        fi = (ul[dim] - ull[dim]) / (uf - ull[dim]);
                        if ((fi <= 0.0) || (fi >= 1.0)) {
                            conv2 = vc * ul[dim];
                        } else {
                            if (fi < b) {
                                if (infacet_ll == 1) conv2 = vc * (a * ul[dim] - c * ull[dim]);
                                else conv2 = vc * ul[dim];
                            }
                            if ((fi >= b) && (fi <= c)) {
                                if (infacet_ll == 1) conv2 = vc * (b * uf + c * ul[dim] - d * ull[dim]);
                                else conv2 = vc * ul[dim];
                            }
                            if (fi > c) conv2 = vc * (c * uf + e * ul[dim]);
                        }
                    }
                } else {
                    if (FLT_EQ(ul[dim], ur[dim])) {
                        conv2 = vc * uf;
                    } else {
                        fi = (uf - ur[dim]) / (ul[dim] - ur[dim]);
                        if ((fi <= 0.0) || (fi >= 1.0)) {
                            conv2 = vc * uf;
                        } else {
                            if (fi < b) conv2 = vc * (a * uf - c * ur[dim]);
                            if ((fi >= b) && (fi <= c)) conv2 = vc * (b * ul[dim] + c * uf - d * ur[dim]);
                            if (fi > c) conv2 = vc * (c * ul[dim] + e * uf);
                        }
                    }
                }
                cc->vc[dim2] = ((conv1 - conv2) / fd[dim]);
            }
        }
        cc->du2dx2[dim2] = du2dx2(fd, dim2, 1.0, cc->ufacet, ul[dim2], ur[dim2]);
    }
}

/* higflow/src/hig-flow-discret.c:541-569 */
static void computational_cell_ve(hfo_solver *s, int flid, const real *fc, const real *fd,
                                  int dim, real *const *dpu)
{
    computational_cell(s, flid, fc, fd, dim, dpu);
    for (int dim2 = 0; dim2 < DIM; dim2++) {
        real Sl, Sr;
        if (dim2 == dim) {
            Sl = center_p_left(s->sdED, fc, fd, dim2, 0.5, s->S[dim][dim2], &s->stn);
            Sr = center_p_right(s->sdED, fc, fd, dim2, 0.5, s->S[dim][dim2], &s->stn);
            s->cc.dSdx[dim2] = dpdx_at(fd, dim2, 0.5, Sl, Sr);
        } else {
            Sl = center_p_2(s->sdED, fc, fd, dim, dim2, 1.0, -1.0, s->S[dim][dim2], &s->stn);
            Sr = center_p_2(s->sdED, fc, fd, dim, dim2, 1.0, +1.0, s->S[dim][dim2], &s->stn);
            s->cc.dSdx[dim2] = dpdx_at(fd, dim2, 1.0, Sl, Sr);
        }
    }
}

/* ------------------------------------------------------------ terms */
/* higflow/src/hig-flow-terms.c:9-305 (single-phase branches) */
static real pressure_term(hfo_solver *s)
{
    return s->par.projtype == INCREMENTAL ? s->cc.dpdx : 0.0;
}

static real source_term(hfo_solver *s) { return s->cc.F; }

static real tensor_term(hfo_solver *s)
{
    real v = 0.0;
    if (s->par.flowtype == VISCOELASTIC)
        for (int d = 0; d < DIM; d++) v += s->cc.dSdx[d];
    return v;
}

static real convective_term(hfo_solver *s, const real *delta)
{
    compcell *cc = &s->cc;
    real v = 0.0;
    switch (cc->convec_type) {
    case CENTRAL:
        for (int d = 0; d < DIM; d++) v += cc->v[d] * cc->dudxc[d];
        break;
    case FIRST_ORDER:
        for (int d = 0; d < DIM; d++) {
            v += 0.5 * (cc->v[d] + fabs(cc->v[d])) * cc->dudxl[d];
            v += 0.5 * (cc->v[d] - fabs(cc->v[d])) * cc->dudxr[d];
        }
        break;
    case SECOND_ORDER:
        if (s->par.secondconvecdiscrtype == CUBISTA) {
            for (int d = 0; d < DIM; d++) v += cc->vc[d];
        } else if (s->par.secondconvecdiscrtype == QUICK) {
            for (int d = 0; d < DIM; d++) {
                v += (cc->v[d] + fabs(cc->v[d])) * 0.1875 * cc->dudxl[d];
                v += (cc->v[d] + fabs(cc->v[d])) * 0.3750 * cc->dudxr[d];
                v -= (cc->v[d] + fabs(cc->v[d])) * 0.0625 * cc->dudxrr[d];
                v += (cc->v[d] - fabs(cc->v[d])) * 0.3750 * cc->dudxl[d];
                v += (cc->v[d] - fabs(cc->v[d])) * 0.1875 * cc->dudxr[d];
                v -= (cc->v[d] - fabs(cc->v[d])) * 0.0625 * cc->dudxll[d];
            }
        } else {
            for (int d = 0; d < DIM; d++) {
                v += 0.5 * (cc->v[d] + fabs(cc->v[d])) * cc->dudxl[d];
                v += 0.5 * (cc->v[d] - fabs(cc->v[d])) * cc->dudxr[d];
                if (cc->v[d] > 0.0) {
                    if (fabs(cc->dudxl[d]) > 1.0e-8)
                        v += 0.5 * cc->v[d] * delta[d] * cc->du2dx2[d] / cc->dudxl[d];
                } else {
                    if (fabs(cc->dudxr[d]) > 1.0e-8)
                        v -= 0.5 * cc->v[d] * delta[d] * cc->du2dx2[d] / cc->dudxr[d];
                }
            }
        }
        break;
    }
    return v;
}

static real diffusive_term(hfo_solver *s)
{
    real v = 0.0;
    for (int d = 0; d < DIM; d++) v += s->cc.du2dx2[d];
    v /= s->par.Re;
    return v;
}

/* --------------------------------------------------- step elements */
/* higflow/src/hig-flow-step.c:613-660 and hig-flow-step-viscoelastic.c:931-980 */
static void explicit_euler_intermediate_velocity(hfo_solver *s, bool ve)
{
    for (int dim = 0; dim < DIM; dim++) {
        hfacetdomain *f = s->sfdu[dim];
        for (int fl = 0; fl < f->nfacets; fl++) {
            Point fc, fd;
            hfo_facet_center(&f->facets[fl], fc);
            hfo_delta(f->facets[fl].c, fd);
            real rhs = 0.0;
            if (!ve) {
                computational_cell(s, fl, fc, fd, dim, s->u);
                rhs -= pressure_term(s);
                rhs -= convective_term(s, fd);
                rhs += diffusive_term(s);
                rhs += source_term(s);
            } else {
                computational_cell_ve(s, fl, fc, fd, dim, s->u);
                rhs += source_term(s);
                rhs -= pressure_term(s);
                rhs += tensor_term(s);
                rhs -= convective_term(s, fd);
                rhs += diffusive_term(s);
            }
            s->ustar[dim][fl] = s->cc.ufacet + s->par.dt * rhs;
        }
    }
}

/* higflow/src/hig-flow-step.c:789-885 and hig-flow-step-viscoelastic.c:1108-1209:
 * explicit convection/pressure/source (and tensor) terms, implicit diffusion
 * (I - dt/Re L) u* = rhs, one system per velocity component. */
static void semi_implicit_euler_intermediate_velocity(hfo_solver *s, bool ve)
{
    for (int dim = 0; dim < DIM; dim++) {
        hfacetdomain *f = s->sfdu[dim];
        hstencil *stn = &s->stn;
        for (int fl = 0; fl < f->nfacets; fl++) {
            Point fc, fd;
            hfo_facet_center(&f->facets[fl], fc);
            hfo_delta(f->facets[fl].c, fd);
            real rhs = 0.0;
            if (!ve) {
                computational_cell(s, fl, fc, fd, dim, s->u);
                rhs += source_term(s);
                rhs -= pressure_term(s);
                rhs -= convective_term(s, fd);
            } else {
                computational_cell_ve(s, fl, fc, fd, dim, s->u);
                rhs += source_term(s);
                rhs -= pressure_term(s);
                rhs += tensor_term(s);
                rhs -= convective_term(s, fd);
            }
            rhs *= s->par.dt;
            rhs += s->cc.ufacet;
            stn_reset(stn);
            stn->rhs = rhs;
            real alpha = 0.0;
            for (int dim2 = 0; dim2 < DIM; dim2++) {
                real w = -s->par.dt / (s->par.Re * fd[dim2] * fd[dim2]);
                alpha -= 2.0 * w;
                Point p;
                memcpy(p, fc, sizeof p);
                p[dim2] = fc[dim2] + fd[dim2];
                hfo_sfd_get_stencil(f, fc, p, w, stn);
                p[dim2] = fc[dim2] - fd[dim2];
                hfo_sfd_get_stencil(f, fc, p, w, stn);
            }
            alpha = 1.0 + alpha;
            hfo_sfd_get_stencil(f, fc, fc, alpha, stn);
            lsys_set_bi(&s->us[dim], fl, stn->rhs);
            lsys_set_row(&s->us[dim], fl, stn->n, stn->ids, stn->vals);
        }
        lsys_solve(&s->us[dim], s->par.rtol, s->par.atol, s->par.maxit);
        memcpy(s->ustar[dim], s->us[dim].x, sizeof(real) * (size_t)f->nfacets);
    }
}

/* higflow/src/hig-flow-step.c:14-45 */
static void remove_pressure_singularity(hfo_solver *s)
{
    if (!s->desingpressure) return;
    real value;
    if (s->par.projtype == INCREMENTAL) value = s->par.pin_pressure - s->p[s->pin];
    else value = s->par.pin_pressure;
    lsys_impose(&s->ps, s->pin, value);
}

/* higflow/src/hig-flow-step.c:48-131 */
static void pressure(hfo_solver *s)
{
    hdomain *sdp = s->sdp;
    remove_pressure_singularity(s);
    hstencil *stn = &s->stn;
    for (int cl = 0; cl < sdp->ncells; cl++) {
        hcell *c = sdp->cells[cl];
        Point cc, cd;
        hfo_center(c, cc);
        hfo_delta(c, cd);
        real sumdudx = 0.0;
        for (int dim = 0; dim < DIM; dim++) {
            int inf;
            real usl = facet_u_left(s->sfdu[dim], cc, cd, dim, 0.5, s->ustar[dim], stn, &inf);
            real usr = facet_u_right(s->sfdu[dim], cc, cd, dim, 0.5, s->ustar[dim], stn, &inf);
            sumdudx += dudxc(cd, dim, 0.5, usl, usl, usr);
        }
        stn_reset(stn);
        stn->rhs = sumdudx / s->par.dt;
        real alpha = 0.0;
        for (int dim = 0; dim < DIM; dim++) {
            real w = 1.0 / (cd[dim] * cd[dim]);
            alpha -= 2.0 * w;
            Point p;
            memcpy(p, cc, sizeof p);
            p[dim] = cc[dim] + cd[dim];
            hfo_sd_get_stencil(sdp, cc, p, w, stn);
            p[dim] = cc[dim] - cd[dim];
            hfo_sd_get_stencil(sdp, cc, p, w, stn);
        }
        hfo_sd_get_stencil(sdp, cc, cc, alpha, stn);
        if (s->desingpressure && cl == s->ps.imposed) continue;
        lsys_set_bi(&s->ps, cl, stn->rhs);
        lsys_set_row(&s->ps, cl, stn->n, stn->ids, stn->vals);
    }
    lsys_solve(&s->ps, s->par.rtol, s->par.atol, s->par.maxit);
    real *dst = (s->par.projtype == INCREMENTAL) ? s->dp : s->p;
    memcpy(dst, s->ps.x, sizeof(real) * (size_t)sdp->ncells);
}

/* higflow/src/hig-flow-step.c:242-308 */
static void final_velocity(hfo_solver *s)
{
    const real *pp = (s->par.projtype == INCREMENTAL) ? s->dp : s->p;
    for (int dim = 0; dim < DIM; dim++) {
        hfacetdomain *f = s->sfdu[dim];
        for (int fl = 0; fl < f->nfacets; fl++) {
            Point fc, fd;
            hfo_facet_center(&f->facets[fl], fc);
            hfo_delta(f->facets[fl].c, fd);
            real pl = center_p_left(s->sdp, fc, fd, dim, 0.5, pp, &s->stn);
            real pr = center_p_right(s->sdp, fc, fd, dim, 0.5, pp, &s->stn);
            real dpdx = dpdx_at(fd, dim, 0.5, pl, pr);
            s->u[dim][fl] = s->ustar[dim][fl] - s->par.dt * dpdx;
        }
    }
}

/* higflow/src/hig-flow-step.c:311-340 */
static void final_pressure(hfo_solver *s)
{
    if (s->par.projtype != INCREMENTAL) return;
    for (int cl = 0; cl < s->sdp->ncells; cl++) s->p[cl] = s->p[cl] + s->dp[cl];
}

/* ---------------------------------------------------- viscoelastic */
/* higflow/src/hig-flow-step-viscoelastic.c:1517-1600 */
static void jacobi(real A[DIM][DIM], real d[DIM], real V[DIM][DIM])
{
    real bb[DIM], z[DIM];
    const int NMAX = 500;
    for (int ip = 0; ip < DIM; ip++) {
        for (int iq = 0; iq < DIM; iq++) V[ip][iq] = 0.0;
        V[ip][ip] = 1.0;
        bb[ip] = A[ip][ip];
        d[ip] = bb[ip];
        z[ip] = 0.0;
    }
    for (int i = 1; i <= NMAX; i++) {
        real tresh, sm = 0.0;
        for (int ip = 0; ip < DIM - 1; ip++)
            for (int iq = ip + 1; iq < DIM; iq++) sm += fabs(A[ip][iq]);
        if (sm == 0.0) return;
        if (i < 4) tresh = 0.2 * sm / (DIM * DIM);
        else tresh = 0.0;
        for (int ip = 0; ip < DIM - 1; ip++)
            for (int iq = ip + 1; iq < DIM; iq++) {
                real g = 100.0 * fabs(A[ip][iq]);
                if ((i > 4) && (fabs(d[ip]) + g == fabs(d[ip])) && (fabs(d[iq]) + g == fabs(d[iq]))) {
                    A[ip][iq] = 0.0;
                } else if (fabs(A[ip][iq]) > tresh) {
                    real t, theta;
                    real h = d[iq] - d[ip];
                    if (fabs(h) + g == fabs(h)) {
                        t = A[ip][iq] / h;
                    } else {
                        theta = 0.5 * h / A[ip][iq];
                        t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
                        if (theta < 0.0) t = -t;
                    }
                    real cc = 1.0 / sqrt(1.0 + t * t);
                    real ss = t * cc;
                    real tau = ss / (1.0 + cc);
                    h = t * A[ip][iq];
                    z[ip] = z[ip] - h;
                    z[iq] = z[iq] + h;
                    d[ip] = d[ip] - h;
                    d[iq] = d[iq] + h;
                    A[ip][iq] = 0.0;
                    for (int j = 0; j <= ip - 1; j++) {
                        g = A[j][ip];
                        h = A[j][iq];
                        A[j][ip] = g - ss * (h + g * tau);
                        A[j][iq] = h + ss * (g - h * tau);
                    }
                    for (int j = ip + 1; j <= iq - 1; j++) {
                        g = A[ip][j];
                        h = A[j][iq];
                        A[ip][j] = g - ss * (h + g * tau);
                        A[j][iq] = h + ss * (g - h * tau);
                    }
                    for (int j = iq + 1; j < DIM; j++) {
                        g = A[ip][j];
                        h = A[iq][j];
                        A[ip][j] = g - ss * (h + g * tau);
                        A[iq][j] = h + ss * (g - h * tau);
                    }
                    for (int j = 0; j < DIM; j++) {
                        g = V[j][ip];
                        h = V[j][iq];
                        V[j][ip] = g - ss * (h + g * tau);
                        V[j][iq] = h + ss * (g - h * tau);
                    }
                }
            }
        for (int ip = 0; ip < DIM; ip++) {
            bb[ip] = bb[ip] + z[ip];
            d[ip] = bb[ip];
            z[ip] = 0.0;
        }
    }
}

/* higflow/src/hig-flow-step-viscoelastic.c:1603-1628: R^T A R and R A R^T */
static void mat_rt_a_r(real A[DIM][DIM], real R[DIM][DIM], real B[DIM][DIM])
{
    for (int i = 0; i < DIM; i++)
        for (int j = 0; j < DIM; j++) {
            B[i][j] = 0.0;
            for (int k = 0; k < DIM; k++)
                for (int l = 0; l < DIM; l++) B[i][j] += R[k][i] * A[k][l] * R[l][j];
        }
}

static void mat_r_a_rt(real A[DIM][DIM], real R[DIM][DIM], real B[DIM][DIM])
{
    for (int i = 0; i < DIM; i++)
        for (int j = 0; j < DIM; j++) {
            B[i][j] = 0.0;
            for (int k = 0; k < DIM; k++)
                for (int l = 0; l < DIM; l++) B[i][j] += R[i][k] * A[k][l] * R[j][l];
        }
}

/* log-conformation transform callbacks of the example drivers
 * (higflow/example2d_Oldroyd/ns-example-2d.c:21-57) */
static real kernel_inverse(const hfo_solver *s, real lambda)
{
    return s->par.kernel_transform == 1 ? exp(lambda) : lambda;
}

static real kernel_jacobian(const hfo_solver *s, real lambda)
{
    if (s->par.kernel_transform != 1) return 1.0;
    real tol = s->par.kernel_tol;
    return (lambda < tol) ? 1.0 / tol : 1.0 / lambda;
}

/* higflow/src/hig-flow-step-generalized-newtonian.c:14-93 */
static void velocity_derivative_tensor(hfo_solver *s)
{
    hdomain *sd = s->sdED;
    for (int cl = 0; cl < sd->ncells; cl++) {
        Point cc, cd;
        hfo_center(sd->cells[cl], cc);
        hfo_delta(sd->cells[cl], cd);
        for (int dim = 0; dim < DIM; dim++)
            for (int dim2 = 0; dim2 < DIM; dim2++) {
                real ul, ur;
                int inf;
                if (dim == dim2) {
                    ul = facet_u_left(s->sfdu[dim], cc, cd, dim, 0.5, s->u[dim], &s->stn, &inf);
                    ur = facet_u_right(s->sfdu[dim], cc, cd, dim, 0.5, s->u[dim], &s->stn, &inf);
                } else {
                    ul = facet_u_4(s->sfdu[dim], cc, cd, dim, dim2, 1.0, -1.0, s->u[dim], &s->stn);
                    ur = facet_u_4(s->sfdu[dim], cc, cd, dim, dim2, 1.0, +1.0, s->u[dim], &s->stn);
                }
                s->D[dim][dim2][cl] = dudxc(cd, dim2, 0.5, ul, ul, ur);
            }
    }
}

/* higflow/src/hig-flow-step.c:1185-1358 */
static real cubista_cell_term(hfo_solver *s, int vdim, const real *kvals, real kc,
                              const real *cc, const real *cd, int dim)
{
    real vbar, vl, vr, kr, krr, kl, kll, fi, conv1 = 0.0, conv2 = 0.0;
    const real a = 1.7500, b = 0.3750, c = 0.7500, d = 0.1250, e = 0.2500;
    int il, ir, ill, irr, inf;
    kl = center_p_22(s->sdED, cc, cd, dim, 1.0, -1.0, kvals, &s->stn, &il);
    kr = center_p_22(s->sdED, cc, cd, dim, 1.0, +1.0, kvals, &s->stn, &ir);
    kll = center_p_22(s->sdED, cc, cd, dim, 2.0, -1.0, kvals, &s->stn, &ill);
    krr = center_p_22(s->sdED, cc, cd, dim, 2.0, +1.0, kvals, &s->stn, &irr);
    vl = facet_u_left(s->sfdu[vdim], cc, cd, dim, 0.5, s->u[vdim], &s->stn, &inf);
    vr = facet_u_right(s->sfdu[vdim], cc, cd, dim, 0.5, s->u[vdim], &s->stn, &inf);
    vbar = vr;
    if (vbar > 0.0) {
        if (FLT_EQ(kr, kl)) {
            conv1 = vbar * kc;
        } else {
            fi = (kc - kl) / (kr - kl);
            if ((fi <= 0.0) || (fi >= 1.0)) {
                conv1 = vbar * kc;
            } else {
                if (fi < b) conv1 = (il == 1) ? vbar * (a * kc - c * kl) : vbar * kc;
                if ((fi >= b) && (fi <= c))
                    conv1 = (il == 1 && ir == 1) ? vbar * (c * kc + b * kr - d * kl) : vbar * kc;
                if (fi > c) conv1 = (ir == 1) ? vbar * (e * kc + c * kr) : vbar * kc;
            }
        }
    } else {
        if (ir == 1 && irr == 1) {
            if (FLT_EQ(kc, krr)) {
                conv1 = vbar * kr;
            } else {
                fi = (kr - krr) / (kc - krr);
                if ((fi <= 0.0) || (fi >= 1.0)) {
                    conv1 = vbar * kr;
                } else {
                    if (fi < b) conv1 = vbar * (a * kr - c * krr);
                    if ((fi >= b) && (fi <= c)) conv1 = vbar * (c * kr + b * kc - d * krr);
                    if (fi > c) conv1 = vbar * (c * kc + e * kr);
                }
            }
        } else if (ir == 1 && irr == 0) {
            if (FLT_EQ(kc, krr)) {
                conv1 = vbar * kr;
            } else {
                fi = (kr - krr) / (kc - krr);
                if ((fi <= 0.0) || (fi >= 1.0)) {
                    conv1 = vbar * kr;
                } else {
                    if (fi <= c) conv1 = vbar * kr;
                    if (fi > c) conv1 = vbar * (c * kc + e * kr);
                }
            }
        } else {
            vbar = vr;
            conv1 = (vbar > 0.0) ? vbar * kc : vbar * kc;
            vbar = vl;
            conv2 = (vbar > 0.0) ? vbar * kl : vbar * kc;
            return ((conv1 - conv2) / cd[dim]);
        }
    }
    vbar = vl;
    if (vbar > 0.0) {
        if (il == 1 && ill == 1) {
            if (FLT_EQ(kc, kll)) {
                conv2 = vbar * kl;
            } else {
                fi = (kl - kll) / (kc - kll);
                if ((fi <= 0.0) || (fi >= 1.0)) {
                    conv2 = vbar * kl;
                } else {
                    if (fi < b) conv2 = vbar * (a * kl - c * kll);
                    if ((fi >= b) && (fi <= c)) conv2 = vbar * (b * kc + c * kl - d * kll);
                    if (fi > c) conv2 = vbar * (c * kc + e * kl);
                }
            }
        } else if (il == 1 && ill == 0) {
            if (FLT_EQ(kc, kll)) {
                conv2 = vbar * kl;
            } else {
                fi = (kl - kll) / (kc - kll);
                if ((fi <= 0.0) || (fi >= 1.0)) {
                    conv2 = vbar * kl;
                } else {
                    if (fi <= c) conv2 = vbar * kl;
                    if (fi > c) conv2 = vbar * (c * kc + e * kl);
                }
            }
        } else {
            vbar = vr;
            conv1 = (vbar > 0.0) ? vbar * kc : vbar * kr;
            vbar = vl;
            conv2 = (vbar > 0.0) ? vbar * kc : vbar * kc;
            return ((conv1 - conv2) / cd[dim]);
        }
    } else {
        if (FLT_EQ(kl, kr)) {
            conv2 = vbar * kc;
        } else {
            fi = (kc - kr) / (kl - kr);
            if ((fi <= 0.0) || (fi >= 1.0)) {
                conv2 = vbar * kc;
            } else {
                if (fi < b) conv2 = (ir == 1) ? vbar * (a * kc - c * kr) : vbar * kc;
                if ((fi >= b) && (fi <= c))
                    conv2 = (il == 1 && ir == 1) ? vbar * (c * kc + b * kl - d * kr) : vbar * kc;
                if (fi > c) conv2 = (il == 1) ? vbar * (e * kc + c * kl) : vbar * kc;
            }
        }
    }
    return ((conv1 - conv2) / cd[dim]);
}

/* higflow/src/hig-flow-step-viscoelastic.c:352-544 (explicit, Oldroyd-B/Giesekus) */
static void explicit_constitutive(hfo_solver *s)
{
    const real De = s->par.De;
    hdomain *sd = s->sdED;
    for (int cl = 0; cl < sd->ncells; cl++) {
        Point cc, cd;
        hfo_center(sd->cells[cl], cc);
        hfo_delta(sd->cells[cl], cd);
        real Du[DIM][DIM], Kn[DIM][DIM], Kc[DIM][DIM];
        for (int i = 0; i < DIM; i++)
            for (int j = 0; j < DIM; j++) {
                Du[i][j] = eval_cell(sd, cc, cc, s->D[i][j], &s->stn);
                (void)eval_cell(sd, cc, cc, s->S[i][j], &s->stn);
                Kn[i][j] = eval_cell(sd, cc, cc, s->K[i][j], &s->stn);
                Kc[i][j] = Kn[i][j];
            }
        real R[DIM][DIM], Kl[DIM], lam[DIM];
        jacobi(Kc, Kl, R);
        for (int k = 0; k < DIM; k++) lam[k] = kernel_inverse(s, Kl[k]);
        real M[DIM][DIM];
        mat_rt_a_r(Du, R, M);
        /* hig_flow_calculate_omega (:1667-1680) */
        real Oa[DIM][DIM], Om[DIM][DIM];
        for (int i = 0; i < DIM - 1; i++) {
            for (int j = i + 1; j < DIM; j++) {
                Oa[i][j] = (M[i][j] * lam[j] + M[j][i] * lam[i]) / (lam[j] - lam[i] + EPSMACH);
                Oa[j][i] = -Oa[i][j];
            }
            Oa[i][i] = 0.0;
        }
        Oa[DIM - 1][DIM - 1] = 0.0;
        mat_r_a_rt(Oa, R, Om);
        real jl[DIM];
        for (int k = 0; k < DIM; k++) jl[k] = kernel_jacobian(s, lam[k]);
        /* hig_flow_calculate_b (:1683-1705) */
        real Ba[DIM][DIM], BB[DIM][DIM];
        for (int i = 0; i < DIM; i++) {
            for (int j = i + 1; j < DIM; j++) {
                Ba[i][j] = 0.0;
                Ba[j][i] = 0.0;
            }
            Ba[i][i] = M[i][i] * lam[i] * jl[i];
        }
        mat_r_a_rt(Ba, R, BB);
        /* hig_flow_calculate_m_oldroyd / _giesekus (:1710-1734) */
        real Ma[DIM][DIM], MM[DIM][DIM];
        for (int i = 0; i < DIM; i++) {
            for (int j = i + 1; j < DIM; j++) {
                Ma[i][j] = 0.0;
                Ma[j][i] = 0.0;
            }
            real aux = 1.0 - lam[i];
            if (s->par.ve_model == GIESEKUS)
                Ma[i][i] = (aux - s->par.giesekus_alpha * aux * aux) * jl[i];
            else
                Ma[i][i] = (1.0 - lam[i]) * jl[i];
        }
        mat_r_a_rt(Ma, R, MM);
        /* hig_flow_kernel_rhs (:1631-1649) */
        real RHS[DIM][DIM];
        {
            real OK[DIM][DIM], KO[DIM][DIM];
            for (int i = 0; i < DIM; i++)
                for (int j = 0; j < DIM; j++) {
                    OK[i][j] = 0.0;
                    KO[i][j] = 0.0;
                    for (int k = 0; k < DIM; k++) {
                        OK[i][j] += Om[i][k] * Kn[k][j];
                        KO[i][j] += Kn[i][k] * Om[k][j];
                    }
                }
            for (int i = 0; i < DIM; i++)
                for (int j = 0; j < DIM; j++)
                    RHS[i][j] = OK[i][j] - KO[i][j] + 2.0 * BB[i][j] + MM[i][j] / De;
        }
        /* hig_flow_velocity_at_center_cell (:1879-1890) */
        real uc[DIM];
        for (int dim = 0; dim < DIM; dim++) {
            int inf;
            real ul = facet_u_left(s->sfdu[dim], cc, cd, dim, 0.5, s->u[dim], &s->stn, &inf);
            real ur = facet_u_right(s->sfdu[dim], cc, cd, dim, 0.5, s->u[dim], &s->stn, &inf);
            uc[dim] = 0.5 * (ul + ur);
        }
        for (int i = 0; i < DIM; i++)
            for (int j = i; j < DIM; j++) {
                real rhs = 0.0;
                if (s->par.ve_convec == CELL_CENTRAL) {
                    /* hig_flow_derivative_kernel_at_center_cell (:1893-1909) */
                    real dKdx[DIM];
                    for (int dim = 0; dim < DIM; dim++) {
                        int ilf, irt;
                        real Kl2 = center_p_22(sd, cc, cd, dim, 1.0, -1.0, s->K[i][j], &s->stn, &ilf);
                        real Kr2 = center_p_22(sd, cc, cd, dim, 1.0, +1.0, s->K[i][j], &s->stn, &irt);
                        if (ilf == 1 && irt == 1) dKdx[dim] = dpdx_at(cd, dim, 1.0, Kl2, Kr2);
                        else if (irt == 1) dKdx[dim] = dpdxr_at(cd, dim, 1.0, Kn[i][j], Kr2);
                        else dKdx[dim] = dpdxl_at(cd, dim, 1.0, Kl2, Kn[i][j]);
                    }
                    for (int dim = 0; dim < DIM; dim++) rhs -= uc[dim] * dKdx[dim];
                } else {
                    for (int dim = 0; dim < DIM; dim++)
                        rhs -= cubista_cell_term(s, dim, s->K[i][j], Kn[i][j], cc, cd, dim);
                }
                rhs += RHS[i][j];
                real kernel = Kn[i][j] + s->par.dt * rhs;
                s->S[i][j][cl] = kernel;
                if (i != j) s->S[j][i][cl] = kernel;
            }
    }
    for (int i = 0; i < DIM; i++)
        for (int j = 0; j < DIM; j++)
            memcpy(s->K[i][j], s->S[i][j], sizeof(real) * (size_t)sd->ncells);
}

/* higflow/src/hig-flow-step-viscoelastic.c:212-347 (Oldroyd-B/Giesekus: fA=a=1, xi=0) */
static void polymeric_tensor(hfo_solver *s)
{
    const real Re = s->par.Re, De = s->par.De, beta = s->par.beta;
    hdomain *sd = s->sdED;
    for (int cl = 0; cl < sd->ncells; cl++) {
        real Kn[DIM][DIM], Du[DIM][DIM];
        for (int i = 0; i < DIM; i++)
            for (int j = 0; j < DIM; j++) {
                Du[i][j] = s->D[i][j][cl];
                Kn[i][j] = s->K[i][j][cl];
            }
        real R[DIM][DIM], lam[DIM];
        jacobi(Kn, lam, R);
        real B[DIM][DIM], A[DIM][DIM];
        for (int i = 0; i < DIM; i++) {
            for (int j = i + 1; j < DIM; j++) {
                B[i][j] = 0.0;
                B[j][i] = 0.0;
            }
            B[i][i] = kernel_inverse(s, lam[i]);
        }
        mat_r_a_rt(B, R, A);
        const real fA = 1.0, a = 1.0, xi = 0.0;
        for (int i = 0; i < DIM; i++) {
            for (int j = 0; j < DIM; j++)
                s->S[i][j][cl] = (1.0 - beta) * (fA * A[i][j]) / (Re * De * (1.0 - xi)) -
                                 (1.0 - beta) * Du[i][j] / Re;
            s->S[i][i][cl] += -a * (1.0 - beta) / (Re * De * (1.0 - xi));
        }
    }
}

/* ------------------------------------------------------------ public */
int hfo_dim(void) { return DIM; }

hfo_solver *hfo_create(void)
{
    hfo_solver *s = calloc(1, sizeof *s);
    s->ids.next = 1;
    stn_init(&s->stn);
    return s;
}

void hfo_destroy(hfo_solver *s)
{
    if (!s) return;
    for (int i = 0; i < s->nspecs; i++) free(s->specs[i].text);
    free(s->specs);
    free(s->timedep);
    if (s->ready) {
        hfo_sd_destroy(s->sdp);
        hfo_sd_destroy(s->sdF);
        hfo_sd_destroy(s->sdED);
        for (int d = 0; d < DIM; d++) {
            hfo_sfd_destroy(s->sfdu[d]);
            hfo_sfd_destroy(s->sfdF[d]);
            free(s->u[d]);
            free(s->ustar[d]);
            free(s->FU[d]);
            for (int e = 0; e < DIM; e++) {
                free(s->K[d][e]);
                free(s->S[d][e]);
                free(s->D[d][e]);
            }
        }
        for (int i = 0; i < s->nspecs; i++) {
            free(s->bc_p[i]->map.lid);
            free(s->bc_p[i]->vals);
            free(s->bc_p[i]->cells);
            free(s->bc_p[i]);
            for (int d = 0; d < DIM; d++) {
                free(s->bc_u[d][i]->map.lid);
                free(s->bc_u[d][i]->vals);
                free(s->bc_u[d][i]->cells);
                free(s->bc_u[d][i]);
            }
            hfo_destroy_tree(s->bctrees_p[i]);
            hfo_destroy_tree(s->bctrees_u[i]);
        }
        free(s->bc_p);
        for (int d = 0; d < DIM; d++) free(s->bc_u[d]);
        free(s->bctrees_p);
        free(s->bctrees_u);
        free(s->p);
        free(s->dp);
        free(s->F);
        lsys_free(&s->ps);
        for (int d = 0; d < DIM; d++) lsys_free(&s->us[d]);
    }
    for (int t = 0; t < s->ntrees; t++) hfo_destroy_tree(s->trees[t]);
    free(s->trees);
    stn_free(&s->stn);
    free(s);
}

int hfo_add_domain_amr(hfo_solver *s, const char *amr_text)
{
    if (s->ready) return -1;
    hcell *root = hfo_read_amr(&s->ids, amr_text);
    if (!root) return -2;
    s->trees = realloc(s->trees, sizeof *s->trees * (size_t)(s->ntrees + 1));
    s->trees[s->ntrees++] = root;
    return s->ntrees - 1;
}

int hfo_add_bc_amr(hfo_solver *s, int userid, const char *amr_text, int ptype,
                   const int *utypes)
{
    if (s->ready) return -1;
    s->specs = realloc(s->specs, sizeof *s->specs * (size_t)(s->nspecs + 1));
    bcspec *b = &s->specs[s->nspecs];
    b->text = strdup(amr_text);
    b->userid = userid;
    b->ptype = ptype;
    for (int d = 0; d < DIM; d++) b->utypes[d] = utypes[d];
    return s->nspecs++;
}

/* Setup order of higflow_initialize_domain_yaml / higflow_create_partitioned_domain
 * (higflow/src/hig-flow-kernel.c:389-402, 486-507, 823-864) followed by the
 * boundary loader (higflow/src/hig-flow-bc.c:27-122, 754-866). */
int hfo_setup(hfo_solver *s, const hfo_params *par)
{
    if (s->ready || s->ntrees == 0) return -1;
    s->par = *par;
    if (s->par.order <= 0) s->par.order = 2;
    int order = s->par.order;
    s->sdp = hfo_sd_create(order);
    s->sdF = hfo_sd_create(order);
    s->sdED = hfo_sd_create(order);
    hdomain *cds[3] = {s->sdp, s->sdF, s->sdED};
    for (int k = 0; k < 3; k++) {
        cds[k]->trees = malloc(sizeof(hcell *) * (size_t)s->ntrees);
        memcpy(cds[k]->trees, s->trees, sizeof(hcell *) * (size_t)s->ntrees);
        cds[k]->ntrees = s->ntrees;
    }
    for (int d = 0; d < DIM; d++) {
        s->sfdu[d] = hfo_sfd_create(d, order, s->trees, s->ntrees);
        hfo_sfd_adjust_facet_ids(s->sfdu[d]);
        s->sfdF[d] = hfo_sfd_create(d, order, s->trees, s->ntrees);
        hfo_sfd_adjust_facet_ids(s->sfdF[d]);
    }
    /* boundary trees are read after the domain, before the maps */
    s->bctrees_p = calloc((size_t)s->nspecs + 1, sizeof(hcell *));
    s->bctrees_u = calloc((size_t)s->nspecs + 1, sizeof(hcell *));
    for (int i = 0; i < s->nspecs; i++) {
        s->bctrees_p[i] = hfo_read_amr(&s->ids, s->specs[i].text);
        s->bctrees_u[i] = hfo_read_amr(&s->ids, s->specs[i].text);
        if (!s->bctrees_p[i] || !s->bctrees_u[i]) return -2;
    }
    int maxid = s->ids.next + 1;
    hfo_sd_build_map(s->sdp, maxid);
    hfo_sd_build_map(s->sdF, maxid);
    hfo_sd_build_map(s->sdED, maxid);
    for (int d = 0; d < DIM; d++) {
        hfo_sfd_compute_sfbi(s->sfdu[d]);
        hfo_sfd_build_map(s->sfdu[d], maxid);
        hfo_sfd_compute_sfbi(s->sfdF[d]);
        hfo_sfd_build_map(s->sfdF[d], maxid);
    }
    s->bc_p = calloc((size_t)s->nspecs + 1, sizeof(hboundary *));
    for (int d = 0; d < DIM; d++) s->bc_u[d] = calloc((size_t)s->nspecs + 1, sizeof(hboundary *));
    s->desingpressure = true;
    for (int i = 0; i < s->nspecs; i++) {
        s->bc_p[i] = hfo_bc_create(s->bctrees_p[i], s->specs[i].ptype, s->specs[i].userid, maxid);
        hfo_sd_add_bc(s->sdp, s->bc_p[i]);
        if (s->specs[i].ptype == HFO_DIRICHLET) s->desingpressure = false;
    }
    for (int d = 0; d < DIM; d++)
        for (int i = 0; i < s->nspecs; i++) {
            s->bc_u[d][i] = hfo_bc_create(s->bctrees_u[i], s->specs[i].utypes[d],
                                          s->specs[i].userid, maxid);
            hfo_sd_add_bc(s->sfdu[d]->cdom, s->bc_u[d][i]);
        }
    int nc = s->sdp->ncells;
    s->p = calloc((size_t)nc, sizeof(real));
    s->dp = calloc((size_t)nc, sizeof(real));
    s->F = calloc((size_t)s->sdF->ncells, sizeof(real));
    for (int d = 0; d < DIM; d++) {
        s->u[d] = calloc((size_t)s->sfdu[d]->nfacets + 1, sizeof(real));
        s->ustar[d] = calloc((size_t)s->sfdu[d]->nfacets + 1, sizeof(real));
        s->FU[d] = calloc((size_t)s->sfdF[d]->nfacets + 1, sizeof(real));
        for (int e = 0; e < DIM; e++) {
            s->K[d][e] = calloc((size_t)s->sdED->ncells, sizeof(real));
            s->S[d][e] = calloc((size_t)s->sdED->ncells, sizeof(real));
            s->D[d][e] = calloc((size_t)s->sdED->ncells, sizeof(real));
        }
    }
    lsys_init(&s->ps, nc);
    for (int d = 0; d < DIM; d++) lsys_init(&s->us[d], s->sfdu[d]->nfacets);
    /* hig-flow-step.c:22-28: the leaf holding the centre of tree 0 */
    {
        Point c0;
        hfo_center(s->trees[0], c0);
        hcell *c2 = hfo_cell_with_point(s->trees[0], c0);
        s->pin = idmap_lookup(&s->sdp->map, c2->cid);
    }
    if (s->par.maxit <= 0) s->par.maxit = 10000;
    s->ready = 1;
    return 0;
}

void hfo_set_time(hfo_solver *s, double t) { s->par.t = t; }
double hfo_get_time(const hfo_solver *s) { return s->par.t; }

static real *field_ptr(hfo_solver *s, int field, int comp, int *n)
{
    if (!s->ready) return NULL;
    int d = comp;
    switch (field) {
    case 0:
        if (d < 0 || d >= DIM) return NULL;
        *n = s->sfdu[d]->nfacets;
        return s->u[d];
    case 1:
        *n = s->sdp->ncells;
        return s->p;
    case 2:
        *n = s->sdp->ncells;
        return s->dp;
    case 3:
        if (d < 0 || d >= DIM) return NULL;
        *n = s->sfdu[d]->nfacets;
        return s->ustar[d];
    case 4:
        *n = s->sdF->ncells;
        return s->F;
    case 5:
        if (d < 0 || d >= DIM) return NULL;
        *n = s->sfdF[d]->nfacets;
        return s->FU[d];
    case 6:
    case 7:
    case 8:
        if (comp < 0 || comp >= DIM * DIM) return NULL;
        *n = s->sdED->ncells;
        if (field == 6) return s->K[comp / DIM][comp % DIM];
        if (field == 7) return s->S[comp / DIM][comp % DIM];
        return s->D[comp / DIM][comp % DIM];
    }
    return NULL;
}

int hfo_field_size(hfo_solver *s, int field, int comp)
{
    int n = 0;
    return field_ptr(s, field, comp, &n) ? n : -1;
}

int hfo_field_centers(hfo_solver *s, int field, int comp, double *xyz)
{
    int n = 0;
    if (!field_ptr(s, field, comp, &n)) return -1;
    for (int i = 0; i < n; i++) {
        Point c;
        if (field == 0 || field == 3) hfo_facet_center(&s->sfdu[comp]->facets[i], c);
        else if (field == 5) hfo_facet_center(&s->sfdF[comp]->facets[i], c);
        else if (field == 4) hfo_center(s->sdF->cells[i], c);
        else if (field >= 6) hfo_center(s->sdED->cells[i], c);
        else hfo_center(s->sdp->cells[i], c);
        for (int k = 0; k < DIM; k++) xyz[i * DIM + k] = c[k];
    }
    return n;
}

int hfo_get_field(hfo_solver *s, int field, int comp, double *out)
{
    int n = 0;
    real *p = field_ptr(s, field, comp, &n);
    if (!p) return -1;
    memcpy(out, p, sizeof(real) * (size_t)n);
    return n;
}

int hfo_set_field(hfo_solver *s, int field, int comp, const double *in)
{
    int n = 0;
    real *p = field_ptr(s, field, comp, &n);
    if (!p) return -1;
    memcpy(p, in, sizeof(real) * (size_t)n);
    return n;
}

static hboundary *bc_of(hfo_solver *s, int kind, int i)
{
    if (!s->ready || i < 0 || i >= s->nspecs) return NULL;
    if (kind == 0) return s->bc_p[i];
    if (kind >= 1 && kind <= DIM) return s->bc_u[kind - 1][i];
    return NULL;
}

int hfo_bc_count(hfo_solver *s, int kind, int i)
{
    hboundary *b = bc_of(s, kind, i);
    return b ? b->nvals : -1;
}

int hfo_bc_centers(hfo_solver *s, int kind, int i, double *xyz)
{
    hboundary *b = bc_of(s, kind, i);
    if (!b) return -1;
    for (int k = 0; k < b->nvals; k++) {
        Point c;
        hfo_center(b->cells[k], c);
        for (int d = 0; d < DIM; d++) xyz[k * DIM + d] = c[d];
    }
    return b->nvals;
}

int hfo_bc_set_values(hfo_solver *s, int kind, int i, const double *vals)
{
    hboundary *b = bc_of(s, kind, i);
    if (!b) return -1;
    memcpy(b->vals, vals, sizeof(real) * (size_t)b->nvals);
    return b->nvals;
}

int hfo_bc_userid(hfo_solver *s, int i)
{
    if (i < 0 || i >= s->nspecs) return -1;
    return s->specs[i].userid;
}

int hfo_set_functions(hfo_solver *s, const hfo_functions *f)
{
    if (!s->ready) return -1;
    if (f) s->fn = *f;
    else memset(&s->fn, 0, sizeof s->fn);
    return 0;
}

int hfo_bc_set_valuetype(hfo_solver *s, int kind, int i, int timedependent)
{
    if (!bc_of(s, kind, i)) return -1;
    if (!s->timedep) s->timedep = calloc((size_t)(DIM + 1) * (size_t)s->nspecs, 1);
    s->timedep[kind * s->nspecs + i] = (char)(timedependent != 0);
    return 0;
}

/* higflow/src/hig-flow-step.c:402-512: time-dependent boundary values.
 * Velocity at t + dt; pressure at t + dt, or its increment over the step
 * for the incremental projection.  The cell and facet source domains of
 * this restatement carry no boundary trees, so :515-606 update nothing. */
static void boundary_refresh(hfo_solver *s)
{
    if (!s->timedep) return;
    const real t = s->par.t, t1 = s->par.t + s->par.dt;
    for (int dim = 0; dim < DIM; dim++) {
        if (!s->fn.get_boundary_velocity) break;
        for (int i = 0; i < s->nspecs; i++) {
            if (!s->timedep[(1 + dim) * s->nspecs + i]) continue;
            hboundary *b = s->bc_u[dim][i];
            for (int k = 0; k < b->nvals; k++) {
                Point c;
                hfo_center(b->cells[k], c);
                b->vals[k] = s->fn.get_boundary_velocity(s->specs[i].userid, c, dim, t1);
            }
        }
    }
    if (!s->fn.get_boundary_pressure) return;
    for (int i = 0; i < s->nspecs; i++) {
        if (!s->timedep[i]) continue;
        hboundary *b = s->bc_p[i];
        for (int k = 0; k < b->nvals; k++) {
            Point c;
            hfo_center(b->cells[k], c);
            real v = s->fn.get_boundary_pressure(s->specs[i].userid, c, t1);
            if (s->par.projtype != NON_INCREMENTAL) v -= s->fn.get_boundary_pressure(s->specs[i].userid, c, t);
            b->vals[k] = v;
        }
    }
}

/* higflow/src/hig-flow-step.c:343-399: cell and facet source terms at t,
 * recomputed every step when the driver registered them */
static void source_refresh(hfo_solver *s)
{
    const real t = s->par.t;
    if (s->fn.get_source_term)
        for (int cl = 0; cl < s->sdF->ncells; cl++) {
            Point c;
            hfo_center(s->sdF->cells[cl], c);
            s->F[cl] = s->fn.get_source_term(c, t);
        }
    if (s->fn.get_facet_source_term)
        for (int dim = 0; dim < DIM; dim++)
            for (int fl = 0; fl < s->sfdF[dim]->nfacets; fl++) {
                Point c;
                hfo_facet_center(&s->sfdF[dim]->facets[fl], c);
                s->FU[dim][fl] = s->fn.get_facet_source_term(c, dim, t);
            }
}

/* higflow/src/hig-flow-step.c:1363-1413 */
int hfo_step(hfo_solver *s)
{
    if (!s->ready) return -1;
    boundary_refresh(s);
    source_refresh(s);
    if (s->par.tempdiscrtype == EXPLICIT_EULER) explicit_euler_intermediate_velocity(s, false);
    else if (s->par.tempdiscrtype == SEMI_IMPLICIT_EULER) semi_implicit_euler_intermediate_velocity(s, false);
    else return -3;
    pressure(s);
    final_velocity(s);
    final_pressure(s);
    s->par.t += s->par.dt;
    return s->ps.converged ? 0 : 1;
}

/* higflow/src/hig-flow-step-viscoelastic.c:2091-2163 */
int hfo_step_viscoelastic(hfo_solver *s)
{
    if (!s->ready) return -1;
    if (s->par.tempdiscrtype != EXPLICIT_EULER && s->par.tempdiscrtype != SEMI_IMPLICIT_EULER) return -3;
    velocity_derivative_tensor(s);
    explicit_constitutive(s);
    polymeric_tensor(s);
    /* the boundary and source refresh follows the tensor updates here
     * (hig-flow-step-viscoelastic.c:2112-2125) */
    boundary_refresh(s);
    source_refresh(s);
    if (s->par.tempdiscrtype == EXPLICIT_EULER) explicit_euler_intermediate_velocity(s, true);
    else semi_implicit_euler_intermediate_velocity(s, true);
    pressure(s);
    final_velocity(s);
    final_pressure(s);
    s->par.t += s->par.dt;
    return s->ps.converged ? 0 : 1;
}

int hfo_advance(hfo_solver *s, int nsteps)
{
    int rc = 0;
    for (int i = 0; i < nsteps && rc >= 0; i++)
        rc = (s->par.flowtype == VISCOELASTIC) ? hfo_step_viscoelastic(s) : hfo_step(s);
    return rc;
}

int hfo_stencil(hfo_solver *s, int domain, int dim, const double *center, const double *x,
                double alpha, int maxn, int *ids, double *w, double *rhs)
{
    if (!s->ready) return -1;
    stn_reset(&s->stn);
    if (domain == 0) hfo_sd_get_stencil(s->sdp, center, x, alpha, &s->stn);
    else if (domain == 1) {
        if (dim < 0 || dim >= DIM) return -1;
        hfo_sfd_get_stencil(s->sfdu[dim], center, x, alpha, &s->stn);
    } else if (domain == 2) hfo_sd_get_stencil(s->sdED, center, x, alpha, &s->stn);
    else return -1;
    int n = s->stn.n < maxn ? s->stn.n : maxn;
    for (int i = 0; i < n; i++) {
        ids[i] = s->stn.ids[i];
        w[i] = s->stn.vals[i];
    }
    *rhs = s->stn.rhs;
    return s->stn.n;
}

int hfo_facet_flag(hfo_solver *s, int dim, const double *x)
{
    hfacet f;
    return hfo_sfd_facet_with_point(s->sfdu[dim], x, &f);
}

int hfo_wls_weights(int dim, int ord, int maxpts, int npts, const double *pts, const double *x,
                    double *w)
{
    if (npts > maxpts || dim > DIM) return -1;
    hwls *wl = hfo_wls_create(dim, ord, maxpts);
    witem *it = calloc((size_t)npts + 1, sizeof *it);
    for (int i = 0; i < npts; i++)
        for (int d = 0; d < dim; d++) it[i].x[d] = pts[i * dim + d];
    hfo_wls_calc(wl, npts, it, x, w);
    free(it);
    hfo_wls_destroy(wl);
    return npts;
}

int hfo_last_solve(hfo_solver *s, int *iters, double *resid)
{
    *iters = s->ps.iters;
    *resid = s->ps.resid;
    return s->ps.converged;
}

int hfo_pinned_cell(hfo_solver *s) { return s->desingpressure ? s->pin : -1; }

/* Test hooks: the cyclic Jacobi eigen-solver and the local (convection-free)
 * constitutive right-hand side of one cell, for closed-form checks. */
int hfo_jacobi(const double *A, double *lam, double *R)
{
    real a[DIM][DIM], v[DIM][DIM], d[DIM];
    for (int i = 0; i < DIM; i++)
        for (int j = 0; j < DIM; j++) a[i][j] = A[i * DIM + j];
    jacobi(a, d, v);
    for (int i = 0; i < DIM; i++) {
        lam[i] = d[i];
        for (int j = 0; j < DIM; j++) R[i * DIM + j] = v[i][j];
    }
    return 0;
}

int hfo_ve_local_rhs(const hfo_params *par, const double *Du, const double *K, double *rhs_out)
{
    hfo_solver tmp;
    memset(&tmp, 0, sizeof tmp);
    tmp.par = *par;
    hfo_solver *s = &tmp;
    real D_[DIM][DIM], Kn[DIM][DIM], Kc[DIM][DIM];
    for (int i = 0; i < DIM; i++)
        for (int j = 0; j < DIM; j++) {
            D_[i][j] = Du[i * DIM + j];
            Kn[i][j] = K[i * DIM + j];
            Kc[i][j] = Kn[i][j];
        }
    real R[DIM][DIM], Kl[DIM], lam[DIM];
    jacobi(Kc, Kl, R);
    for (int k = 0; k < DIM; k++) lam[k] = kernel_inverse(s, Kl[k]);
    real M[DIM][DIM];
    mat_rt_a_r(D_, R, M);
    real Oa[DIM][DIM], Om[DIM][DIM];
    for (int i = 0; i < DIM - 1; i++) {
        for (int j = i + 1; j < DIM; j++) {
            Oa[i][j] = (M[i][j] * lam[j] + M[j][i] * lam[i]) / (lam[j] - lam[i] + EPSMACH);
            Oa[j][i] = -Oa[i][j];
        }
        Oa[i][i] = 0.0;
    }
    Oa[DIM - 1][DIM - 1] = 0.0;
    mat_r_a_rt(Oa, R, Om);
    real jl[DIM], Ba[DIM][DIM], BB[DIM][DIM], Ma[DIM][DIM], MM[DIM][DIM];
    for (int k = 0; k < DIM; k++) jl[k] = kernel_jacobian(s, lam[k]);
    for (int i = 0; i < DIM; i++) {
        for (int j = 0; j < DIM; j++) {
            Ba[i][j] = 0.0;
            Ma[i][j] = 0.0;
        }
        Ba[i][i] = M[i][i] * lam[i] * jl[i];
        real aux = 1.0 - lam[i];
        Ma[i][i] = (s->par.ve_model == GIESEKUS) ? (aux - s->par.giesekus_alpha * aux * aux) * jl[i]
                                                 : (1.0 - lam[i]) * jl[i];
    }
    mat_r_a_rt(Ba, R, BB);
    mat_r_a_rt(Ma, R, MM);
    for (int i = 0; i < DIM; i++)
        for (int j = 0; j < DIM; j++) {
            real ok = 0.0, ko = 0.0;
            for (int k = 0; k < DIM; k++) {
                ok += Om[i][k] * Kn[k][j];
                ko += Kn[i][k] * Om[k][j];
            }
            rhs_out[i * DIM + j] = ok - ko + 2.0 * BB[i][j] + MM[i][j] / s->par.De;
        }
    return 0;
}

/*
 * CPU oracle for the HiG-Flow projection step -- TEST INFRASTRUCTURE ONLY.
 * Public C API used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg (through ctypes).  The product library never links this.
 *
 * Mirrors the reference driver flow (higflow/example2d_Newt/ns-example-2d.c):
 *   create -> add domain .amr -> add boundary .amr + types -> setup(params)
 *   -> set boundary values / initial fields -> step (t += dt) ...
 */
#ifndef HFO_H
#define HFO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfo_solver hfo_solver;

/* Mirrors higflow_parameters / higflow_controllers / ve_parameters
 * (higflow/src/hig-flow-kernel.h:108-214, 425-482) for the fields the
 * projection step reads.  Enum values are the reference's. */
typedef struct hfo_params {
    double Re, dt, t;
    int projtype;              /* NON_INCREMENTAL=0, INCREMENTAL=1 */
    int flowtype;              /* NEWTONIAN=0, VISCOELASTIC=3 */
    int tempdiscrtype;         /* EXPLICIT_EULER=0, SEMI_IMPLICIT_EULER=3 */
    int convecdiscrtype;       /* CENTRAL=0, FIRST_ORDER=1, SECOND_ORDER=2 */
    int secondconvecdiscrtype; /* MCU=0, CUBISTA=1, QUICK=2 */
    /* viscoelastic */
    int ve_model;              /* OLDROYD_B=0, GIESEKUS=1 */
    int ve_convec;             /* CELL_CENTRAL=0, CELL_CUBISTA=1 */
    int kernel_transform;      /* 0 identity (conformation), 1 log-conformation */
    double De, beta, kernel_tol, giesekus_alpha;
    /* linear solvers (PETSc KSP options of the reference run lines) */
    double rtol, atol;
    int maxit;
    /* exact pressure at the desingularised cell: get_pressure(x, t) */
    double pin_pressure;
    int order;                 /* interpolation order, 2 in every driver */
} hfo_params;

int hfo_dim(void);
hfo_solver *hfo_create(void);
void hfo_destroy(hfo_solver *s);
int hfo_add_domain_amr(hfo_solver *s, const char *amr_text);
/* bc types: DIRICHLET=0, NEUMANN=1; utypes has DIM entries */
int hfo_add_bc_amr(hfo_solver *s, int userid, const char *amr_text, int ptype,
                   const int *utypes);
int hfo_setup(hfo_solver *s, const hfo_params *par);
void hfo_set_time(hfo_solver *s, double t);
double hfo_get_time(const hfo_solver *s);

/* fields: 0 u[comp], 1 p, 2 deltap, 3 ustar[comp], 4 F, 5 FU[comp],
 *         6 kernel[comp=i*DIM+j], 7 S[i*DIM+j], 8 Du[i*DIM+j] */
int hfo_field_size(hfo_solver *s, int field, int comp);
int hfo_field_centers(hfo_solver *s, int field, int comp, double *xyz);
int hfo_get_field(hfo_solver *s, int field, int comp, double *out);
int hfo_set_field(hfo_solver *s, int field, int comp, const double *in);

/* boundary values: kind 0 pressure, 1+d velocity component d */
int hfo_bc_count(hfo_solver *s, int kind, int bcindex);
int hfo_bc_centers(hfo_solver *s, int kind, int bcindex, double *xyz);
int hfo_bc_set_values(hfo_solver *s, int kind, int bcindex, const double *vals);
int hfo_bc_userid(hfo_solver *s, int bcindex);

/* higflow_functions (higflow/src/hig-flow-kernel.h:216-233), the callbacks
 * the step re-evaluates: boundary values of boundaries marked time-dependent
 * (hfo_bc_set_valuetype; kind 0 pressure, 1+d velocity component d) and,
 * when set, the cell and facet source terms.  Point arguments are real[DIM]. */
typedef struct hfo_functions {
    double (*get_boundary_pressure)(int id, double *center, double t);
    double (*get_boundary_velocity)(int id, double *center, int dim, double t);
    double (*get_source_term)(double *center, double t);
    double (*get_facet_source_term)(double *center, int dim, double t);
} hfo_functions;
int hfo_set_functions(hfo_solver *s, const hfo_functions *f);
int hfo_bc_set_valuetype(hfo_solver *s, int kind, int bcindex, int timedependent);

/* one projection step, then t += dt (ns-example-2d.c:240-246) */
int hfo_step(hfo_solver *s);
int hfo_step_viscoelastic(hfo_solver *s);
int hfo_advance(hfo_solver *s, int nsteps);

/* point stencil query: domain 0 = pressure cells, 1 = velocity facets of
 * component dim, 2 = viscoelastic cell domain.  Returns the element count. */
int hfo_stencil(hfo_solver *s, int domain, int dim, const double *center, const double *x,
                double alpha, int maxn, int *ids, double *w, double *rhs);
int hfo_facet_flag(hfo_solver *s, int dim, const double *x);

/* raw WLS: dim/ord/maxpts as wls_create, points row-major npts x dim */
int hfo_wls_weights(int dim, int ord, int maxpts, int npts, const double *pts,
                    const double *x, double *w);

/* last pressure solve */
int hfo_last_solve(hfo_solver *s, int *iters, double *resid);
int hfo_pinned_cell(hfo_solver *s);

/* test hooks: cyclic Jacobi (row-major DIM x DIM) and the convection-free
 * constitutive right-hand side dK/dt of one cell */
int hfo_jacobi(const double *A, double *lam, double *R);
int hfo_ve_local_rhs(const hfo_params *par, const double *Du, const double *K, double *rhs);

#ifdef __cplusplus
}
#endif
#endif

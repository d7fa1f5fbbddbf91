/*
 * higflow_hip: the HiG-Flow projection time step and the HiG-Tree linear
 * solver plugin on one AMD Instinct MI355X (gfx950).
 *
 * Plain C ABI: opaque handles, plain pointers and sizes.  Every entry point
 * returns a negative value on error (hfh_last_error() holds the message) and
 * never aborts the caller.
 *
 * Two surfaces, each mirroring the reference interface it replaces:
 *
 *  1. Solver-plugin engine  (hfh_slv_*): one function per member of
 *     HiG-Tree's _solver_vtable (higtree/src/solver.h:42-60) plus the
 *     generic slv_* wrappers that carry behaviour (slv_impose_value,
 *     solver.c:142-149; imposed-line skipping in slv_set_bi/add_bi,
 *     solver.c:99-113; post-solve SolverIterations/SolverResidual/converged,
 *     solver.h:80-82).  Global ids are first_gid-based, as in slv_create
 *     (solver.h:90).
 *
 *  2. Projection step  (hfh_*): the HiG-Flow step entry points
 *     higflow_solver_step (higflow/src/hig-flow-step.h:82) and
 *     higflow_solver_step_viscoelastic (hig-flow-step-viscoelastic.h:36)
 *     on a device-resident copy of the solver state, built from the same
 *     inputs the reference driver reads (domain and boundary .amr trees,
 *     boundary types and values, higflow_parameters/controllers), plus the
 *     HiG-Tree point-stencil query sd_get_stencil/sfd_get_stencil
 *     (higtree/src/domain.h:317-335) on the library's own mesh.
 */
#ifndef HIGFLOW_HIP_H
#define HIGFLOW_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ common */
/* Message of the last failed call on this thread ("" if none). */
const char *hfh_last_error(void);
/* Number of visible HIP devices (0 when none; never fails). */
int hfh_device_count(void);
/* Library version: major*10000 + minor*100 + patch. */
int hfh_version(void);

/* ------------------------------------------------- 1. solver-plugin engine */
typedef struct hfh_slv hfh_slv;

/* slv_create (higtree/src/solver.c:239-331) for the HIP engine */
hfh_slv *hfh_slv_create(int first_gid, int size);
/* _solver_vtable.set_maxnonzeros (solver.h:43) */
int hfh_slv_set_maxnonzeros(hfh_slv *s, int maxnonzeros);
/* _solver_vtable.set_Ai (solver.h:44): replaces global row i */
int hfh_slv_set_Ai(hfh_slv *s, int i, int n, const int *cols, const double *vals);
/* slv_set_bi / slv_add_bi (solver.c:99-113): ignored on the imposed line */
int hfh_slv_set_bi(hfh_slv *s, int i, double v);
int hfh_slv_add_bi(hfh_slv *s, int i, double v);
/* _solver_vtable.set_xi / set_x / get_xi / get_x / get_x_scatter (solver.h:47-51) */
int hfh_slv_set_xi(hfh_slv *s, int i, double v);
int hfh_slv_set_x(hfh_slv *s, const double *x);
double hfh_slv_get_xi(hfh_slv *s, int i);
int hfh_slv_get_x(hfh_slv *s, double *x);
int hfh_slv_get_x_scatter(hfh_slv *s, int n, const int *idx, double *x);
/* slv_impose_value (solver.c:142-149): identity row, column kept */
int hfh_slv_impose_value(hfh_slv *s, int line, double value);
/* _solver_vtable.assemble_matrix / assemble_rhs / assemble_output (solver.h:52-54) */
int hfh_slv_assemble_matrix(hfh_slv *s);
int hfh_slv_assemble_rhs(hfh_slv *s);
int hfh_slv_assemble_output(hfh_slv *s);
/* _solver_vtable.solve (solver.h:55): BiCGStab (right Jacobi) on the device */
int hfh_slv_solve(hfh_slv *s);
/* _solver_vtable.destroy (solver.h:56) */
void hfh_slv_destroy(hfh_slv *s);
/* tolerances as slv_set_rel_tol/slv_set_abs_tol/slv_set_maxiteration */
int hfh_slv_set_tolerances(hfh_slv *s, double rtol, double atol, int maxit);
/* post-solve fields (solver.h:80-82) */
int hfh_slv_iterations(hfh_slv *s);
double hfh_slv_residual(hfh_slv *s);
int hfh_slv_converged(hfh_slv *s);

/* ------------------------------------------------------ 2. projection step */
typedef struct hfh_solver hfh_solver;

/* higflow_parameters / higflow_controllers / ve_parameters fields the step
 * reads (higflow/src/hig-flow-kernel.h:108-214, 425-482); enum values are
 * the reference's. */
typedef struct hfh_params {
    double Re, dt, t;
    int projtype;              /* NON_INCREMENTAL=0, INCREMENTAL=1 */
    int flowtype;              /* NEWTONIAN=0, VISCOELASTIC=3 */
    int tempdiscrtype;         /* EXPLICIT_EULER=0 */
    int convecdiscrtype;       /* CENTRAL=0, FIRST_ORDER=1, SECOND_ORDER=2 */
    int secondconvecdiscrtype; /* MCU=0, CUBISTA=1, QUICK=2 */
    int ve_model;              /* OLDROYD_B=0, GIESEKUS=1 */
    int ve_convec;             /* CELL_CENTRAL=0, CELL_CUBISTA=1 */
    int kernel_transform;      /* 0 conformation, 1 log-conformation */
    double De, beta, kernel_tol, giesekus_alpha;
    double rtol, atol;         /* Poisson tolerances (-ksp_rtol/-ksp_atol) */
    int maxit;
    double pin_pressure;       /* get_pressure() at the desingularised cell */
    int order;                 /* interpolation order (2 in every driver) */
    int precision;             /* 0 FP64, 1 mixed FP32-inner/FP64 refinement */
    double inner_rtol;         /* mixed: inner FP32 tolerance */
} hfh_params;

/* Stage ids for hfh_timing / hfh_stage_bytes */
enum { HFH_ST_INTERP = 0, HFH_ST_MOMENTUM, HFH_ST_DIV, HFH_ST_SOLVE, HFH_ST_PROJECT, HFH_ST_VE, HFH_NSTAGE };

hfh_solver *hfh_create(int dim);
void hfh_destroy(hfh_solver *s);
/* higio_read_from_amr text of a domain tree (higtree/src/higtree-io.c:269-293) */
int hfh_add_domain_amr(hfh_solver *s, const char *amr_text);
/* a boundary tree with its pressure type and per-component velocity types
 * (DIRICHLET=0, NEUMANN=1), as one bcN entry of <prefix>.bc.yaml */
int hfh_add_bc_amr(hfh_solver *s, int userid, const char *amr_text, int ptype, const int *utypes);
/* builds domains, boundary objects, the device operators and state */
int hfh_setup(hfh_solver *s, const hfh_params *par);

/* boundary values at boundary-cell centres (sb_set_value): kind 0 pressure,
 * 1+d velocity component d */
int hfh_bc_count(hfh_solver *s, int kind, int bcindex);
int hfh_bc_centers(hfh_solver *s, int kind, int bcindex, double *xyz);
int hfh_bc_set_values(hfh_solver *s, int kind, int bcindex, const double *vals);

/* The driver callbacks of higflow_functions (higflow/src/hig-flow-kernel.h:
 * 216-233) that higflow_solver_step re-evaluates every step; Point arguments
 * are real[DIM], so the driver's own functions fit these pointers.  On each
 * step, on the host, before the intermediate velocity (after the tensor
 * updates in the viscoelastic step):
 *  - boundaries marked time-dependent with hfh_bc_set_valuetype take
 *    get_boundary_velocity(userid, centre, dim, t + dt) and
 *    get_boundary_pressure(userid, centre, t + dt) (minus its value at t for
 *    the incremental projection), hig-flow-step.c:402-512;
 *  - F = get_source_term(cell centre, t) and FU = get_facet_source_term(
 *    facet centre, dim, t) on every cell and facet when the pointer is
 *    non-null, hig-flow-step.c:343-399.
 * A null member is never called; f = NULL clears them all.  Not for a batch. */
typedef struct hfh_functions {
    double (*get_boundary_pressure)(int id, double *center, double t);
    double (*get_boundary_velocity)(int id, double *center, int dim, double t);
    double (*get_source_term)(double *center, double t);
    double (*get_facet_source_term)(double *center, int dim, double t);
} hfh_functions;
int hfh_set_functions(hfh_solver *s, const hfh_functions *f);
/* valuetype of one boundary for one kind (0 pressure, 1+d velocity
 * component d): 0 fixed (sb_set_value by the driver only), 1 time-dependent
 * (bc_valuetype timedependent, hig-flow-bc.c:12-26) */
int hfh_bc_set_valuetype(hfh_solver *s, int kind, int bcindex, int timedependent);

/* fields: 0 u[comp], 1 p, 2 deltap, 3 ustar[comp], 4 F, 5 FU[comp],
 *         6 kernel[i*DIM+j], 7 S[i*DIM+j], 8 Du[i*DIM+j],
 *         9 polymeric stress T[i*DIM+j] = S + (1 - beta) Du / Re (read only,
 *           computed on request; the tensor of the reference's Tmin/Tmax
 *           line, hig-flow-step-viscoelastic.c:319-335)
 * in the reference's local-id order (cells: depth-first leaves; facets:
 * the facet iterator of domain.c:1834-1894). */
int hfh_field_size(hfh_solver *s, int field, int comp);
int hfh_field_centers(hfh_solver *s, int field, int comp, double *xyz);
int hfh_get_field(hfh_solver *s, int field, int comp, double *out);
int hfh_set_field(hfh_solver *s, int field, int comp, const double *in);
/* min and max of a field on the device, from +-1e16 (for u[comp]: the
 * Vmin/Vmax that higflow_final_velocity prints, hig-flow-step.c:250-306);
 * returns the field size.  With HFH_VERBOSE=1 every step prints that line,
 * and a viscoelastic step first prints the Tmin/Tmax line of field 9 for
 * each i >= j. */
int hfh_field_range(hfh_solver *s, int field, int comp, double *vmin, double *vmax);
/* No counterpart in the reference: hfh_field_range for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops
 * that log the velocity range every step).  min and max of the device array
 * x[0, n), from +-1e16 as hfh_field_range starts them, written to the device
 * array minmax[2]; enqueued on `stream` (a hipStream_t; NULL is the null
 * stream), with no host synchronisation.  It calls no hipMalloc/hipFree:
 * its partials live in a fixed scratch array of the library (64 KiB per
 * device, held until the process ends), so n may not exceed
 * HFH_RANGE_NOALLOC_MAX.  Calls that overlap on one device (other streams
 * or threads) share that scratch and race.  Returns 0, or <0 with
 * hfh_last_error() set. */
#define HFH_RANGE_NOALLOC_MAX 33554432L
int hfh_range_noalloc(const double *x, long n, double *minmax, void *stream);
/* lowpoint[DIM] then highpoint[DIM] of every pressure cell (hig_cell,
 * higtree.h:59-75), in local-id order: the geometry the reference's VTK
 * writer prints (hig-flow-io.c:233-265); returns the cell count */
int hfh_cell_bounds(hfh_solver *s, double *lohi);

/* higflow_solver_step (hig-flow-step.c:1363-1413); then t += dt */
int hfh_solver_step(hfh_solver *s);
/* higflow_solver_step_viscoelastic (hig-flow-step-viscoelastic.c:2091-2163) */
int hfh_solver_step_viscoelastic(hfh_solver *s);
int hfh_set_time(hfh_solver *s, double t);
double hfh_get_time(hfh_solver *s);
int hfh_pinned_cell(hfh_solver *s);
int hfh_last_solve(hfh_solver *s, int *iters, double *resid, int *outer);
/* the last solve of a system (0 pressure, 1+d semi-implicit velocity
 * component d), as the solver's SolverIterations / SolverResidual /
 * converged (higtree/src/solver.h:80-82): info = {iterations, refinement
 * steps, mixed form's FP64 use (0 FP32 refinement, 1 FP32 failed and the
 * solve finished in FP64, 2 FP64 chosen as the faster form), true-residual
 * restarts}; returns 1 converged, 0 not converged, <0 error */
int hfh_solve_stats(hfh_solver *s, int which, int *info, double *resid);

/* sd_get_stencil / sfd_get_stencil on the library's host mesh: domain 0
 * pressure cells, 1 velocity facets of component dim, 2 tensor cells.
 * Returns the element count (ids/w filled up to maxn), rhs in *rhs. */
int hfh_stencil(hfh_solver *s, int domain, int dim, const double *center, const double *x, double alpha,
                int maxn, int *ids, double *w, double *rhs);

/* device timing (hipEvents on the library stream), per stage: total ms and
 * number of timed launches; algorithmic bytes of one stage launch */
int hfh_timing(hfh_solver *s, double *ms, long *calls);
int hfh_reset_timing(hfh_solver *s);
long hfh_stage_bytes(hfh_solver *s, int stage);
/* per-stage event timing inside hfh_solver_step* (off by default: it adds a
 * host synchronisation per stage) */
int hfh_set_timing(hfh_solver *s, int on);
/* launches one step kernel `reps` times between hipEvents on the library
 * stream: 0 pressure SpMV (+dot), 1 momentum, 2 momentum interpolation rows,
 * 3 divergence rows, 4 projection rows, 5 constitutive, 6 / 7 the fused
 * pressure BiCGStab SpMV launch in FP64 / FP32, 8 a plain copy of 64 MiB,
 * 9 / 10 the fused FP64 BiCGStab launch A (update + SpMV) / launch B (SpMV)
 * of the semi-implicit velocity block system (the pressure system's when
 * the step has none), 11 / 12 the fused PCG SpMV launch of that velocity
 * system in FP64 / FP32, 13 the fine-level prolongation + post-smoothing
 * launch of that system's FP32 V-cycle (AMG-PCG velocity systems), 14 the
 * same launch of the pressure system's FP32 V-cycle, 15 the fine-level
 * residual launch of the pressure system's FP64 V-cycle (CSR rows).
 * Average ms per launch and the algorithmic bytes of one launch; -1 when
 * the step has no such launch.  Call between steps. */
int hfh_probe(hfh_solver *s, int which, int reps, double *avg_ms, long *bytes);
int hfh_sync(hfh_solver *s);
/* the assembled systems as CSR: which 0 pressure (the matrix slv_set_Ai
 * receives in higflow_pressure, hig-flow-step.c:48-131, with the pinned
 * unknown eliminated and rows negated), 1+d the semi-implicit velocity
 * system of component d (hig-flow-step.c:789-885).  Call with null rp/ci/v
 * for the sizes; *sym is 1 when the matrix is symmetric. */
int hfh_operator_csr(hfh_solver *s, int which, int *n, int *nnz, int *sym, int *rp, int *ci, double *v);

/* No counterpart in the reference: an inspection aid for callers and tests
 * that check set-up determinism.  A 64-bit FNV-1a hash of the recorded host
 * tables (interpolation groups, slot tables, pressure and Helmholtz
 * systems) as held now; equal for set-ups that recorded the same step
 * (e.g. on different HFH_SETUP_THREADS).  A device solver relabels some
 * tables at upload, so compare solvers of the same kind. */
int hfh_tables_digest(hfh_solver *s, unsigned long long *out);

/* ------------------------------------------------- 3. batched parameter sweeps
 * No counterpart in the reference, which runs a sweep as one driver process
 * per simulation (higflow/example2d_Oldroyd/ns-example-2d.c:465-471 is one
 * run's time loop).  Independent simulations on meshes of different sizes
 * advance together: one launch per step stage for all of them, and their
 * linear systems solved block-diagonally with one stopping test per
 * simulation (the same tolerances and iterates as a run of its own).
 *
 * hfh_create_host: a solver that builds its mesh, stencils and operator
 * tables on the host only (hfh_add_*, hfh_setup and hfh_bc_set_values as
 * usual; stepping fails).  hfh_create_batch takes n such members, all with
 * the same time step, schemes and solver settings (Re, De and beta may
 * differ), merges them into one device solver and frees the member handles
 * (members[i] is set to NULL).  Step the batch with hfh_solver_step /
 * hfh_solver_step_viscoelastic; hfh_probe, hfh_timing and hfh_sync work as
 * for one run; fields, boundary values and solve statistics go per member
 * through hfh_batch_*. */
hfh_solver *hfh_create_host(int dim);
hfh_solver *hfh_create_batch(int dim, hfh_solver **members, int n);
int hfh_batch_size(hfh_solver *s);
int hfh_batch_field_size(hfh_solver *s, int member, int field, int comp);
int hfh_batch_get_field(hfh_solver *s, int member, int field, int comp, double *out);
int hfh_batch_set_field(hfh_solver *s, int member, int field, int comp, const double *in);
int hfh_batch_bc_set_values(hfh_solver *s, int member, int kind, int bcindex, const double *vals);
/* as hfh_solve_stats, for one member's system */
int hfh_batch_solve_stats(hfh_solver *s, int member, int which, int *info, double *resid);

#ifdef __cplusplus
}
#endif
#endif

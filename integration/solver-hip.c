/*
 * SOLVER_HIP: a HiG-Tree linear-solver engine backed by libhigflow_hip.
 *
 * Drop this file into higtree/src/ next to solver-petsc.c and register it
 * with three edits:
 *   - "hip" in engine_names and SOLVER_HIP in SolverEngine
 *     (higtree/src/solver.h:14-39, same position in both lists);
 *   - case SOLVER_HIP: s = _slv_create_hip(first_gid, size); in slv_create
 *     (higtree/src/solver.c:280-310).
 * HIGTREE_SOLVER_ENGINE=hip then selects it (solver.c:249-259).
 *
 * Each vtable member (solver.h:42-60) forwards to the hfh_slv_* entry of the
 * same name.  Like the PETSc engine (_solver_impl_petsc, solver-petsc.c:61-62)
 * the generic `solver` is the first member of the engine's struct, and
 * slv_destroy frees that struct after the engine's destroy (solver.c:228-236),
 * so destroy releases only the device solver.  The imposed row of
 * slv_impose_value (solver.c:142-149) and the imposed-line skip of
 * slv_set_bi/slv_add_bi (:99-113) stay in the generic layer.
 *
 * Link with -L<repo>/higflow_amd -lhigflow_hip -Wl,-rpath,<repo>/higflow_amd.
 */
#include <stdio.h>
#include <stdlib.h>

#include "solver.h"
#include "higflow_hip.h"

typedef struct {
    solver s;
    hfh_slv *h;
} _solver_impl_hip;

#define H(cs) (((_solver_impl_hip *)(cs))->h)

static void hip_set_maxnonzeros(solver *cs) { hfh_slv_set_maxnonzeros(H(cs), cs->MaxNonZeros); }
static void hip_set_Ai(solver *cs, int i, int n, const int *j, const real *v) { hfh_slv_set_Ai(H(cs), i, n, j, v); }
static void hip_set_bi(solver *cs, int i, real v) { hfh_slv_set_bi(H(cs), i, v); }
static void hip_add_bi(solver *cs, int i, real v) { hfh_slv_add_bi(H(cs), i, v); }
static void hip_set_xi(solver *cs, int i, real v) { hfh_slv_set_xi(H(cs), i, v); }
static void hip_set_x(solver *cs, const real *x) { hfh_slv_set_x(H(cs), x); }
static real hip_get_xi(solver *cs, int i) { return hfh_slv_get_xi(H(cs), i); }
static void hip_get_x(solver *cs, real *x) { hfh_slv_get_x(H(cs), x); }
static void hip_get_x_scatter(solver *cs, int n, const int *idx, real *x) { hfh_slv_get_x_scatter(H(cs), n, idx, x); }
static void hip_assemble_matrix(solver *cs) { hfh_slv_assemble_matrix(H(cs)); }
static void hip_assemble_rhs(solver *cs) { hfh_slv_assemble_rhs(H(cs)); }
static void hip_assemble_output(solver *cs) { hfh_slv_assemble_output(H(cs)); }

static void hip_solve(solver *cs)
{
    /* tolerances as the caller left them in the generic struct */
    hfh_slv_set_tolerances(H(cs), cs->relative_tolerance, cs->absolute_tolerance, cs->MaxIteration);
    if (hfh_slv_solve(H(cs)) < 0) fprintf(stderr, "SOLVER_HIP: %s\n", hfh_last_error());
    cs->SolverIterations = hfh_slv_iterations(H(cs));
    cs->SolverResidual = hfh_slv_residual(H(cs));
    cs->converged = hfh_slv_converged(H(cs)) == 1;
}

static void hip_destroy(solver *cs)
{
    hfh_slv_destroy(H(cs));
    H(cs) = NULL;
}

static const _solver_vtable _hip_vtable = {
    .set_maxnonzeros = hip_set_maxnonzeros,
    .set_Ai = hip_set_Ai,
    .set_bi = hip_set_bi,
    .add_bi = hip_add_bi,
    .set_xi = hip_set_xi,
    .set_x = hip_set_x,
    .get_xi = hip_get_xi,
    .get_x = hip_get_x,
    .get_x_scatter = hip_get_x_scatter,
    .assemble_matrix = hip_assemble_matrix,
    .assemble_rhs = hip_assemble_rhs,
    .assemble_output = hip_assemble_output,
    .solve = hip_solve,
    .destroy = hip_destroy,
};

solver *_slv_create_hip(int first_gid, int size);

solver *_slv_create_hip(int first_gid, int size)
{
    _solver_impl_hip *impl = calloc(1, sizeof *impl);
    if (!impl) return NULL;
    impl->h = hfh_slv_create(first_gid, size);
    if (!impl->h) {
        fprintf(stderr, "SOLVER_HIP: %s\n", hfh_last_error());
        free(impl);
        return NULL;
    }
    impl->s.vtable = &_hip_vtable;
    /* slv_create leaves these to the engine (its generic defaults sit
     * behind `if(!SOLVER_PETSC)`, solver.c:318-322): the reference run
     * lines' -ksp_rtol/-ksp_atol 1e-10 */
    impl->s.MaxIteration = 10000;
    impl->s.relative_tolerance = 1e-10;
    impl->s.absolute_tolerance = 1e-10;
    return &impl->s;
}

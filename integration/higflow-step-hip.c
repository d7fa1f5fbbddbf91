/*
 * Binds HiG-Flow's step entry points (higflow_solver_step,
 * hig-flow-step.h:82; higflow_solver_step_viscoelastic,
 * hig-flow-step-viscoelastic.h:36) to libhigflow_hip without changing the
 * higflow_solver layout (hig-flow-kernel.h:1142-1194): attached solvers are
 * kept in a small table keyed by the higflow_solver pointer.
 *
 * What the binding reads from the reference's own objects:
 *   ns->par (Re, dt, t), ns->contr (projection, flow type, time and
 *   convection schemes), ns->ed.ve.{par,contr} and the kernel callbacks
 *   (viscoelastic flows), the boundary types, user ids and value types of
 *   ns->sdp and ns->sfdu[d] (domain.h:29-38, 84-124), the solver tolerances
 *   of ns->slvp (solver.h:62-86), and ns->func (hig-flow-kernel.h:216-233):
 *   boundary values at set-up as hig-flow-bc.c:29-122 sets them, the
 *   time-dependent refresh and the source terms every step.
 * Fields move as dp->values[0 .. local_count) (pdomain.h:97-123, 258-261),
 * whose order is the library's field order.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "higflow-step-hip.h"
#include "higflow_hip.h"

#define MAXBIND 16

static struct {
    higflow_solver *ns;
    hfh_solver *h;
    int ve;
} g_bind[MAXBIND];

static hfh_solver *lookup(higflow_solver *ns, int *ve)
{
    for (int i = 0; i < MAXBIND; i++)
        if (g_bind[i].ns == ns && ns) {
            if (ve) *ve = g_bind[i].ve;
            return g_bind[i].h;
        }
    return NULL;
}

static char *read_text(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *s = malloc((size_t)n + 1);
    if (s && fread(s, 1, (size_t)n, f) != (size_t)n) {
        free(s);
        s = NULL;
    }
    if (s) s[n] = '\0';
    fclose(f);
    return s;
}

/* boundary of user id `id` in a domain: its type and value type */
static int find_bc(sim_domain *sd, int id, int *type, int *timedep)
{
    for (int i = 0; i < sd->numdirichlet_bcs; i++)
        if (sd->dirichlet_bcs[i]->userid == id) {
            *type = DIRICHLET;
            *timedep = sd->dirichlet_bcs[i]->valuetype == timedependent;
            return 0;
        }
    for (int i = 0; i < sd->numneumann_bcs; i++)
        if (sd->neumann_bcs[i]->userid == id) {
            *type = NEUMANN;
            *timedep = sd->neumann_bcs[i]->valuetype == timedependent;
            return 0;
        }
    return -1;
}

/* the kernel transform the driver registered: 0 identity (conformation
 * form), 1 logarithm (log-conformation), -1 another one (no GPU path) */
static int kernel_transform(higflow_solver *ns)
{
    const real tol = ns->ed.ve.par.kernel_tol;
    const real x[2] = {0.5, 2.0};
    int ident = 1, lg = 1;
    for (int k = 0; k < 2; k++) {
        real v = ns->ed.ve.get_kernel(0, x[k], tol);
        real w = ns->ed.ve.get_kernel_inverse(0, v, tol);
        ident = ident && v == x[k] && w == x[k];
        lg = lg && fabs(v - log(x[k])) <= 1e-15 && fabs(w - x[k]) <= 1e-14;
    }
    return ident ? 0 : lg ? 1 : -1;
}

struct fieldmap {
    int field, comp;
    distributed_property *dp;
};

static int field_list(higflow_solver *ns, int ve, struct fieldmap *out)
{
    int n = 0;
    for (int d = 0; d < DIM; d++) out[n++] = (struct fieldmap){0, d, ns->dpu[d]};
    out[n++] = (struct fieldmap){1, 0, ns->dpp};
    out[n++] = (struct fieldmap){2, 0, ns->ddeltap};
    for (int d = 0; d < DIM; d++) out[n++] = (struct fieldmap){3, d, ns->dpustar[d]};
    out[n++] = (struct fieldmap){4, 0, ns->dpF};
    for (int d = 0; d < DIM; d++) out[n++] = (struct fieldmap){5, d, ns->dpFU[d]};
    if (ve)
        for (int i = 0; i < DIM; i++)
            for (int j = 0; j < DIM; j++) {
                out[n++] = (struct fieldmap){6, i * DIM + j, ns->ed.ve.dpKernel[i][j]};
                out[n++] = (struct fieldmap){7, i * DIM + j, ns->ed.ve.dpS[i][j]};
                out[n++] = (struct fieldmap){8, i * DIM + j, ns->ed.ve.dpD[i][j]};
            }
    return n;
}

static int move_fields(higflow_solver *ns, int to_device)
{
    int ve;
    hfh_solver *h = lookup(ns, &ve);
    if (!h) return -1;
    struct fieldmap fm[6 + 4 * DIM + 3 * DIM * DIM];
    int n = field_list(ns, ve, fm);
    for (int k = 0; k < n; k++) {
        distributed_property *dp = fm[k].dp;
        if (!dp) continue;
        int size = hfh_field_size(h, fm[k].field, fm[k].comp);
        if (size < 0 || size != dp->pdata->local_count) return -1;
        int rc = to_device ? hfh_set_field(h, fm[k].field, fm[k].comp, dp->values)
                           : hfh_get_field(h, fm[k].field, fm[k].comp, dp->values);
        if (rc < 0) return -1;
    }
    return 0;
}

int higflow_hip_push(higflow_solver *ns) { return move_fields(ns, 1); }
int higflow_hip_pull(higflow_solver *ns) { return move_fields(ns, 0); }

static void set_boundary_values(higflow_solver *ns, hfh_solver *h, int nbcs, const int bc_id[])
{
    const real t = ns->par.t, dt = ns->par.dt;
    for (int b = 0; b < nbcs; b++)
        for (int kind = 0; kind <= DIM; kind++) {
            int n = hfh_bc_count(h, kind, b);
            if (n <= 0) continue;
            double *xyz = malloc(sizeof(double) * (size_t)n * DIM);
            double *val = malloc(sizeof(double) * (size_t)n);
            hfh_bc_centers(h, kind, b, xyz);
            for (int k = 0; k < n; k++) {
                real *c = xyz + (size_t)k * DIM;
                if (kind == 0) {
                    /* hig-flow-bc.c:54-66: pressure at t + dt, or its
                     * increment for the incremental projection */
                    val[k] = ns->func.get_boundary_pressure(bc_id[b], c, t + dt);
                    if (ns->contr.projtype == INCREMENTAL) val[k] -= ns->func.get_boundary_pressure(bc_id[b], c, t);
                } else {
                    /* hig-flow-bc.c:103-109: velocity at t */
                    val[k] = ns->func.get_boundary_velocity(bc_id[b], c, kind - 1, t);
                }
            }
            hfh_bc_set_values(h, kind, b, val);
            free(xyz);
            free(val);
        }
}

int higflow_hip_attach(higflow_solver *ns, int ndomains, const char *const domain_amr[], int nbcs,
                       const int bc_id[], const char *const bc_amr[])
{
    int slot = -1;
    for (int i = 0; i < MAXBIND && slot < 0; i++)
        if (!g_bind[i].ns) slot = i;
    if (slot < 0 || lookup(ns, NULL)) return -1;
    const int ve = ns->contr.flowtype == VISCOELASTIC;
    if (ns->contr.flowtype != NEWTONIAN && !ve) return -1;
    if (ns->contr.tempdiscrtype != EXPLICIT_EULER && ns->contr.tempdiscrtype != SEMI_IMPLICIT_EULER) return -1;
    if (ns->contr.spatialdiscrtype != ORDER2) return -1;

    hfh_params p;
    memset(&p, 0, sizeof p);
    p.Re = ns->par.Re;
    p.dt = ns->par.dt;
    p.t = ns->par.t;
    p.projtype = ns->contr.projtype;
    p.flowtype = ns->contr.flowtype;
    p.tempdiscrtype = ns->contr.tempdiscrtype;
    p.convecdiscrtype = ns->contr.convecdiscrtype;
    p.secondconvecdiscrtype = ns->contr.secondconvecdiscrtype;
    p.order = 2;
    if (ve) {
        if (ns->ed.ve.contr.model != OLDROYD_B && ns->ed.ve.contr.model != GIESEKUS) return -1;
        if (ns->ed.ve.contr.discrtype != EXPLICIT) return -1;
        p.ve_model = ns->ed.ve.contr.model;
        p.ve_convec = ns->ed.ve.contr.convecdiscrtype;
        p.kernel_transform = kernel_transform(ns);
        if (p.kernel_transform < 0) return -1;
        p.De = ns->ed.ve.par.De;
        p.beta = ns->ed.ve.par.beta;
        p.kernel_tol = ns->ed.ve.par.kernel_tol;
        p.giesekus_alpha = ns->ed.ve.par.alpha;
    }
    p.rtol = ns->slvp ? ns->slvp->relative_tolerance : 1e-10;
    p.atol = ns->slvp ? ns->slvp->absolute_tolerance : 1e-10;
    p.maxit = ns->slvp ? ns->slvp->MaxIteration : 10000;

    hfh_solver *h = hfh_create(DIM);
    if (!h) return -1;
    int rc = 0;
    for (int i = 0; i < ndomains && rc >= 0; i++) {
        char *txt = read_text(domain_amr[i]);
        rc = txt ? hfh_add_domain_amr(h, txt) : -1;
        free(txt);
    }
    sim_domain *sdp = psd_get_local_domain(ns->psdp);
    int timedep[DIM + 1][64] = {{0}};
    for (int b = 0; b < nbcs && rc >= 0; b++) {
        int ptype, utypes[3] = {0, 0, 0};
        if (b >= 64 || find_bc(sdp, bc_id[b], &ptype, &timedep[0][b]) < 0) {
            rc = -1;
            break;
        }
        for (int d = 0; d < DIM && rc >= 0; d++) {
            sim_facet_domain *sfd = psfd_get_local_domain(ns->psfdu[d]);
            rc = find_bc(sfd->cdom, bc_id[b], &utypes[d], &timedep[1 + d][b]);
        }
        char *txt = rc >= 0 ? read_text(bc_amr[b]) : NULL;
        rc = txt ? hfh_add_bc_amr(h, bc_id[b], txt, ptype, utypes) : -1;
        free(txt);
    }
    if (rc >= 0) {
        /* the pinned cell of remove_pressure_singularity (hig-flow-step.c:
         * 14-45): the leaf of the first tree that holds the tree's centre;
         * its get_pressure value at the first step's t + dt */
        Point c0;
        hig_get_center(sd_get_higtree(sdp, 0), c0);
        hig_cell *leaf = hig_get_cell_with_point(sd_get_higtree(sdp, 0), c0);
        hig_get_center(leaf, c0);
        p.pin_pressure = ns->func.get_pressure ? ns->func.get_pressure(c0, ns->par.t + ns->par.dt) : 0.0;
        rc = hfh_setup(h, &p);
    }
    if (rc >= 0) {
        set_boundary_values(ns, h, nbcs, bc_id);
        for (int kind = 0; kind <= DIM; kind++)
            for (int b = 0; b < nbcs; b++)
                if (timedep[kind][b]) hfh_bc_set_valuetype(h, kind, b, 1);
        hfh_functions f = {
            .get_boundary_pressure = ns->func.get_boundary_pressure,
            .get_boundary_velocity = ns->func.get_boundary_velocity,
            .get_source_term = ns->func.get_source_term,
            .get_facet_source_term = ns->func.get_facet_source_term,
        };
        rc = hfh_set_functions(h, &f);
    }
    if (rc < 0) {
        hfh_destroy(h);
        return -1;
    }
    g_bind[slot].ns = ns;
    g_bind[slot].h = h;
    g_bind[slot].ve = ve;
    if (higflow_hip_push(ns) < 0) {
        higflow_hip_detach(ns);
        return -1;
    }
    return 0;
}

void higflow_hip_detach(higflow_solver *ns)
{
    for (int i = 0; i < MAXBIND; i++)
        if (g_bind[i].ns == ns && ns) {
            hfh_destroy(g_bind[i].h);
            g_bind[i].ns = NULL;
            g_bind[i].h = NULL;
        }
}

int higflow_hip_step(higflow_solver *ns)
{
    int ve;
    hfh_solver *h = lookup(ns, &ve);
    if (!h) return -1;
    /* the driver advances ns->par.t after the step (ns-example-2d.c:244-246) */
    hfh_set_time(h, ns->par.t);
    int rc = ve ? hfh_solver_step_viscoelastic(h) : hfh_solver_step(h);
    if (rc < 0) {
        fprintf(stderr, "higflow_hip_step: %s\n", hfh_last_error());
        return rc;
    }
    /* host fields for the driver's print/save test that follows the time
     * update (ns-example-2d.c:248-262) */
    const real t1 = ns->par.t + ns->par.dt;
    if (t1 >= ns->par.tp || t1 >= ns->par.ts) higflow_hip_pull(ns);
    return 0;
}
